"""The reference's tracer API, backed by libpach.so on one MI355X.

Names, argument meanings and call order follow Pachyderm_Acoustic_Universal:

    scene = Polygon_Scene(20, 50, 1013.25, 0, False, True)      # Classes_Scenes.cs:442
    scene.Construct(model, materials, scatterings, transmissions)  # :447-630
    scene.partition()                                            # :709-755
    src = GeodesicSource([120] * 8, (2, 3, 1.5), 0, False)       # Classes_Sources.cs:278
    rec = Receiver_Bank(points, src, scene, 1000, 2000.0)        # Classes_Receivers.cs:139
    rt = SplitRayTracer(src, rec, scene, 2.0 * scene.Sound_speed(0), [0, 7], -1, 100000, None)
    rt.Begin()                                                   # TaskAll:209
    while rt.ThreadState() == "Running": ...                     # TaskAll:246
    rt.Combine_ThreadLocal_Results()                             # TaskAll:768
    etc = rt.GetReceiver.GetEnergyHistogram(4, 0, 0)             # Classes_Receivers.cs:432

Every trace runs through the C ABI; there is no CPU path.
"""
from __future__ import annotations

import time

import numpy as np

from . import _abi as A

Running, Stopped = "Running", "Stopped"


# ---------------------------------------------------------------- materials
class Basic_Material:
    """Octave absorption coefficients (Classes_Materials.cs:60-300)."""

    def __init__(self, ABS):
        a = np.asarray(ABS, dtype=np.float64)
        if a.size == 24:  # third octaves averaged per octave (:128-139)
            a = a.reshape(8, 3).mean(1)
        if a.size != 8:
            raise ValueError("Basic_Material needs 8 octave (or 24 third-octave) coefficients")
        self.Abs = a

    def Coefficient_A_Broad(self, Octave=None):
        return self.Abs.copy() if Octave is None else float(self.Abs[Octave])


class Lambert_Scattering:
    """Octave scattering coefficients (Classes_Materials.cs:4463-4485)."""

    def __init__(self, Scattering):
        s = np.asarray(Scattering, dtype=np.float64)
        if s.size != 8:
            raise ValueError("Lambert_Scattering needs 8 octave coefficients")
        self.Scat = s

    def Coefficient(self, octave=None):
        return self.Scat.copy() if octave is None else float(self.Scat[octave])


# ---------------------------------------------------------------- scene
class Polygon_Scene:
    """Polygon scene with a uniform medium (Classes_Scenes.cs:414-1031).

    Pa is the pressure in hPa, as Uniform_Medium expects it (Classes_Medium_Props.cs:288-289)."""

    def __init__(self, Temp, hr, Pa, Air_Choice=0, EdgeCorrection=False, IsAcoustic=True):
        self.TempC, self.hr, self.Pa = float(Temp), float(hr), float(Pa)
        self.atten, self.c, self.rho = A.uniform_medium(Temp, hr, Pa, Air_Choice, EdgeCorrection)
        self._tris = None
        self.grid = (0, 0, 0)
        self.Valid = False

    def Construct(self, Model, Mat, Scat, Trans=None):
        """Model[object][polygon] = list of (x, y, z) vertices; one material,
        scattering and transmission entry per object, copied to each polygon
        (:484-487).  Polygons are fan-triangulated; triangles keep the polygon's
        index as Surf_ID."""
        if len(Mat) != len(Model) and len(Scat) != len(Model):
            raise ValueError("The number of material codes must match the number of objects in the model")
        tris, surf, alpha, scat, trans, areas, objects = [], [], [], [], [], [], []
        pid = 0
        for i, obj in enumerate(Model):
            tr = np.zeros(8) if Trans is None else np.asarray(Trans[i], dtype=np.float64)
            for poly in obj:
                P = np.asarray(poly, dtype=np.float64)
                if P.shape[0] < 3:
                    raise ValueError("polygon with fewer than 3 vertices")
                area = 0.0
                for k in range(1, P.shape[0] - 1):
                    tris.append([P[0], P[k], P[k + 1]])
                    surf.append(pid)
                    area += 0.5 * np.linalg.norm(np.cross(P[k] - P[0], P[k + 1] - P[0]))
                alpha.append(Mat[i].Coefficient_A_Broad())
                scat.append(Scat[i].Coefficient())
                trans.append(tr)
                areas.append(area)
                objects.append(i)
                pid += 1
        self._tris = np.asarray(tris)
        self._surf = np.asarray(surf, dtype=np.int32)
        self._alpha, self._scat, self._trans = map(np.asarray, (alpha, scat, trans))
        self._area = np.asarray(areas)
        self._object = np.asarray(objects, dtype=np.int32)  # Object_ID (:498-502)
        self.Valid = True

    @classmethod
    def from_room(cls, room, Temp=20.0, hr=50.0, Pa=1013.25, Air_Choice=0):
        """Scene of a rooms.Room (already triangulated, materials per polygon)."""
        s = cls(Temp, hr, Pa, Air_Choice)
        s._tris = np.asarray(room.verts, dtype=np.float64)
        s._surf = np.asarray(room.surf, dtype=np.int32)
        s._alpha, s._scat, s._trans = (np.asarray(x, dtype=np.float64) for x in (room.alpha, room.scat, room.trans))
        s._area = np.zeros(room.n_surfaces)
        s._object = np.arange(room.n_surfaces, dtype=np.int32)
        for t, p in zip(s._tris, s._surf):
            s._area[p] += 0.5 * np.linalg.norm(np.cross(t[1] - t[0], t[2] - t[0]))
        s.Valid = True
        return s

    def partition(self, SP_Param=None, Max_Polys=None):
        """Voxel grid of SP_Param^3 voxels (None: sized from the triangle count)."""
        if SP_Param is not None:
            self.grid = tuple(SP_Param) if hasattr(SP_Param, "__len__") else (SP_Param,) * 3

    def Count(self):
        return int(self._alpha.shape[0])

    @property
    def ObjectCount(self):
        """Objects (Model[i] of Construct), not polygons (:967-973)."""
        return int(self._object.max()) + 1

    def ObjectID(self, i):
        return int(self._object[i])

    def SurfaceArea(self, x):
        return float(self._area[x])

    def Sound_speed(self, arg=0):
        return self.c

    def Attenuation(self, arg=0):
        return self.atten.copy()

    def Rho(self, arg=0):
        return self.rho

    def Rho_C(self, arg=0):
        return self.rho * self.c


# ---------------------------------------------------------------- sources
class Source:
    """Omni point source; power W = 1e-12 10^(0.1 SWL) per band (Classes_Sources.cs:50-104)."""

    dir_mode = A.DIR_GEODESIC
    type_name = "Geodesic"  # Source.Type() (Classes_Sources.cs:220-223, :282)

    def __init__(self, power_In_db, Source, ID=0, Third_Octave=False):
        if Third_Octave:
            raise NotImplementedError("third-octave sources are not traced by this library")
        self.SPL = np.asarray(power_In_db, dtype=np.float64)
        self.SourcePower = 1e-12 * 10 ** (0.1 * self.SPL)
        self.Center = np.asarray(Source, dtype=np.float64)
        self.S_ID = ID
        self.dirs = None

    @property
    def Origin(self):
        return self.Center

    @property
    def SoundPower(self):
        return self.SourcePower

    def SWL(self, octave=None):
        if octave is None:
            return self.SPL.copy()
        if octave < 8:
            return float(self.SPL[octave])
        return 10 * np.log10(np.sum(10 ** (0.1 * self.SPL)))

    def Revolution_Period(self):
        return 80


class GeodesicSource(Source):
    """Stratified over the 80 faces of an order-1 geodesic sphere (Classes_Sources.cs:272-404)."""

    dir_mode = A.DIR_GEODESIC


class RandomSource(Source):
    """theta, phi ~ U(0, 2 pi) (Classes_Sources.cs:238-266)."""

    dir_mode = A.DIR_RANDOM
    type_name = "PseudoRandom"


class SuppliedDirectionSource(Source):
    """Source ray i leaves along dirs[i % n] (the deterministic specular mode's seeded list)."""

    dir_mode = A.DIR_SUPPLIED
    type_name = "Supplied"

    def __init__(self, power_In_db, Source, dirs, ID=0):
        super().__init__(power_In_db, Source, ID)
        self.dirs = np.asarray(dirs, dtype=np.float64).reshape(-1, 3)


# ---------------------------------------------------------------- receivers
class Receiver_Bank:
    """Spherical receivers with directional histograms (Classes_Receivers.cs:39-693).

    radius: the reference fixes 1 m (:751); other radii keep the expected
    energy unchanged (SizeMod = 1 / (pi r^3))."""

    def __init__(self, Pt, SrcPT, Sc, SampleRate_in, COTime_in, Type="Stationary", third_Oct=False,
                 radius=1.0):
        if Type != "Stationary" or third_Oct:
            raise NotImplementedError("only stationary octave-band receivers are traced")
        self.Origins = np.asarray(Pt, dtype=np.float64).reshape(-1, 3)
        self.SampleRate = int(SampleRate_in)
        self.CutOffTime = float(COTime_in)
        self.SampleCT = int(np.floor(COTime_in * SampleRate_in / 1000))
        self.radius = float(radius)
        self.SWL = SrcPT.SWL()
        self.scene = Sc
        self.planes = np.zeros((A.NPLANES, len(self.Origins), 8, self.SampleCT))
        self.scale_n = 1.0

    @property
    def Count(self):
        return len(self.Origins)

    def Origin(self, i):
        return self.Origins[i]

    def Scale(self, Ray_CT):
        """Receiver_Bank.Scale (Classes_Receivers.cs:676-679, :1774-1786)."""
        n = float(Ray_CT)
        p = A.PLANES["pressure"]
        pressure = self.planes[p] / np.sqrt(n)
        self.planes /= n
        self.planes[p] = pressure

    def _filter_channels(self, rec):
        """Directional_Histogram.Create_Filter's envelopes (Classes_Receivers.cs:
        1911-1936): pressure, then per axis P neg / E and P pos / E (0 where
        E = 0) -- [7][8][SampleCT]."""
        E, P = self.planes[0, rec], self.planes[1, rec]
        zero = E == 0
        safe = np.where(zero, 1.0, E)
        ch = [P]
        for d in range(3):
            ch.append(np.where(zero, 0.0, P * self.planes[5 + d, rec] / safe))
            ch.append(np.where(zero, 0.0, P * self.planes[2 + d, rec] / safe))
        return np.array(ch)

    def _filters(self, recs, swl, device=0):
        chans = np.concatenate([self._filter_channels(r) for r in recs])
        with A.Tracer(device) as tr:
            F = tr.etc_to_filter(chans, swl, self.SampleRate)
        return F[:, :-4096].reshape(len(recs), 7, -1)

    def Create_Filter(self, SWL=None, Rec_ID=None, dim=None, VB=None):
        """Auralisation filters (Receiver_Bank.Create_Filter, Classes_Receivers.cs:
        626-663; Directional_Histogram.Create_Filter :1911-1995), computed on the
        GPU (pach_etc_to_filter) from the scaled histograms.

        Create_Filter(): every receiver with SWL 120 dB; sets F [R][n] (omni)
        and Fdir [R][6][n] (per axis: the P neg / E, then the P pos / E filter).
        Create_Filter(SWL, Rec_ID): that receiver's 7 filters (omni, then per
        axis neg, pos).  Create_Filter(SWL, Rec_ID, dim): one of them (dim 0
        omni; 2d + 1 and 2d + 2 the neg / pos filters of axis d, zero where
        E = 0; the reference's own single-direction path fails there)."""
        if SWL is None or not hasattr(SWL, "__len__"):
            F = self._filters(range(self.Count), np.full(8, 120.0))
            self.F, self.Fdir = F[:, 0], F[:, 1:]
            return None
        F = self._filters([int(Rec_ID)], np.asarray(SWL, dtype=np.float64))[0]
        if dim is None:
            return F
        return F[int(dim) if 0 <= int(dim) <= 6 else 0]

    def GetEnergyHistogram(self, Octave, Delay_ms, Rec_Index):
        """Energy histogram padded by the delay (:432-438); Octave 8 sums the bands."""
        E = self.planes[0, Rec_Index]
        h = E.sum(0) if Octave >= 8 else E[Octave]
        pad = int(np.floor(Delay_ms / 1000.0 * self.SampleRate))
        return np.concatenate([np.zeros(pad), h])

    def GetPressureHistogram(self, Octave, Rec_Index):
        return self.planes[1, Rec_Index, Octave].copy()

    def Directions_Pos(self, Octave, Rec_Index):
        return self.planes[2:5, Rec_Index, Octave].copy()

    def Directions_Neg(self, Octave, Rec_Index):
        return self.planes[5:8, Rec_Index, Octave].copy()

    def Reflection_Analysis_Data(self, Rec_Index):
        """Spherical_Receiver.Reflection_Analysis_Data of one receiver
        (Classes_Receivers.cs:714, :857-875): [3][8][SampleCT], [0] specular
        arrivals never scattered, [1] arrivals of scattered rays, [2] specular
        arrivals of rays scattered before."""
        return self.planes[8:11, Rec_Index].copy()


# ---------------------------------------------------------------- mapping receivers
class PachMapReceiver(Receiver_Bank):
    """Mapping receivers (Pach_Map_Receiver.cs:34-1690) over a mesh.

    Map_Mesh: (vertices [V][3], polygons: lists of 3 or 4 vertex indices).
    Receivers sit at the polygon centres offset by the unit polygon normal,
    or (RecOnVertex) at the vertices offset by increment along the vertex
    normal, or with Z_displacement at centre / vertex + increment / 2 in z
    (Partition :99-181); spheres of diameter Increment, recording at 1000 Hz
    (the rate the reference fixes), times relative to each receiver's direct
    arrival.  A SplitRayTracer with this bank as RecMain traces into the map
    (pach_map_receivers_set).  With Time1Pt each receiver keeps one sample
    per band (SampleCT 1, :93): Histogram_1Pt adds every deposit's energy
    there whatever its time (Classes_Receivers.cs:1393-1396)."""

    def __init__(self, Map_Mesh, Src_Pt, SampleRate_in, Increment_in, Sc, Cutoff_time, Time1Pt=False,
                 Z_displacementIn=False, DirectionalIn=True, RecOnVertex=False, Offset_Mesh=False):
        V = np.asarray(Map_Mesh[0], dtype=np.float64).reshape(-1, 3)
        polys = [list(p) for p in Map_Mesh[1]]
        self.increment = float(Increment_in)
        self.Directional = bool(DirectionalIn)
        self.Z_Displacement = bool(Z_displacementIn)
        self.Rec_Vertex = bool(RecOnVertex)
        self.Mesh_Offset = bool(Offset_Mesh)
        self.Time1Pt = bool(Time1Pt)
        self.Src = np.asarray(Src_Pt.Origin, dtype=np.float64)
        self.CutOffTime = float(Cutoff_time)
        self.SampleRate = 1000
        self.SampleCT = 1 if self.Time1Pt else int(np.floor(self.CutOffTime * self.SampleRate / 1000))
        self.SWL = Src_Pt.SWL()
        self.scene = Sc
        self.Radius = self.increment * 0.5
        normals = []
        for p in polys:
            n = np.cross(V[p[1]] - V[p[0]], V[p[2]] - V[p[0]])
            normals.append(n / np.linalg.norm(n))
        normals = np.array(normals)
        if self.Rec_Vertex:
            vn = np.zeros_like(V)
            for p, n in zip(polys, normals):
                for v in p:
                    vn[v] += n
            vn /= np.maximum(np.linalg.norm(vn, axis=1, keepdims=True), 1e-300)
            pts = V + (np.array([0, 0, self.increment * .5]) if self.Z_Displacement else vn * self.increment)
        else:
            centers = np.array([V[p].mean(0) for p in polys])
            pts = centers + (np.array([0, 0, self.increment * .5]) if self.Z_Displacement else normals)
        self.Origins = pts
        self.Map_Mesh = (V, polys)
        self.Direct_Time = np.linalg.norm(self.Src - pts, axis=1) / Sc.Sound_speed(0)
        self.planes = np.zeros((8, len(pts), 8, self.SampleCT))
        self.delay_ms = 0.0

    @property
    def Count(self):
        return len(self.Origins)

    @staticmethod
    def Get_SPL_Map(Rec_List, T_Bounds, Octave, SrcID, Coherent_Superposition=False, ZeroAtDirect=False):
        """Incoherent SPL map (Pach_Map_Receiver.cs:410-540, the
        Coherent_Superposition = false branch): the energy of samples
        [floor(T_Bounds[1]), floor(T_Bounds[0])) of each receiver's histogram,
        summed over the sources with their arrival offsets, each source's
        histogram weighted by Pressure_Intensity(1) = sqrt(405) as written."""
        if Coherent_Superposition:
            raise NotImplementedError("the coherent map needs the pressure filters of each receiver")
        SrcID = list(SrcID)
        first = Rec_List[0]
        out = np.zeros(first.Count)
        t_max, t_min = int(np.floor(T_Bounds[0])), int(np.floor(T_Bounds[1]))
        w = np.sqrt(1.0 * 405.0)
        for i in range(first.Count):
            hist = np.zeros(first.SampleCT)
            zero = 0
            if ZeroAtDirect:
                zero = int(min(Rec_List[s].Direct_Time[i] for s in SrcID) * first.SampleRate)
            for s in SrcID:
                temp = Rec_List[s].GetEnergyHistogram(Octave, Rec_List[s].delay_ms, i)
                arrival = int(Rec_List[s].Direct_Time[i] * first.SampleRate) - zero
                for j in range(temp.size):
                    t = j + arrival
                    if hist.size == 1:
                        hist[0] += temp[j] * w
                    if t >= hist.size:
                        break
                    hist[t] += temp[j] * w
            E = hist[t_min:t_max].sum()
            out[i] = AcousticalMath.SPL_Intensity(E)
        return out

    @staticmethod
    def Get_Clarity_Map(Rec_List, C_Cutoff, Octave, SrcID, Coherent_Superposition=False):
        """Incoherent clarity map (:925-991): Clarity of each receiver's
        energy-time curve from the start of its histogram."""
        if Coherent_Superposition:
            raise NotImplementedError("the coherent map needs the pressure filters of each receiver")
        first = Rec_List[0]
        out = np.zeros(first.Count)
        for i in range(first.Count):
            etc = sum(Rec_List[s].GetEnergyHistogram(Octave, 0.0, i) for s in SrcID)
            out[i] = AcousticalMath.Clarity(etc, first.SampleRate, C_Cutoff / 1000.0, 0.0, False)
        return out


# ---------------------------------------------------------------- tracer
class Simulation_Type:
    """Abstract simulation contract (Classes_Simulation_Type.cs:27-35)."""

    def Sim_Type(self):
        raise NotImplementedError

    def Begin(self):
        raise NotImplementedError

    def ProgressMsg(self):
        raise NotImplementedError

    def ThreadState(self):
        raise NotImplementedError

    def Combine_ThreadLocal_Results(self):
        raise NotImplementedError


class SplitRayTracer(Simulation_Type):
    """Stochastic split ray tracer (TaskAll:33-782) on one MI355X.

    cutoffTime is a path length in metres (the ctor divides it by the speed of
    sound, TaskAll:83).  rayCountIn < 1 selects the convergence mode
    (TaskAll:111-176, :288-329): revolutions of 80 source rays until the
    checks of the farthest visible and the farthest occluded receiver pass
    (rayCount < 0: Minimum_Convergence_Check, 0: SpecMin_Convergence_Check).
    Extra keyword arguments:
    precision ("fp32" traversal, or "fp64"), mode ("stochastic" or "specular"),
    seed (the reference seeds from the clock; None does the same), device."""

    def __init__(self, sourceIn, receiverIn, roomIn, cutoffTime, octaveRange, isOrderIn, rayCountIn,
                 Vis_Feedback=None, *, precision="fp32", mode="stochastic", seed=None, device=0,
                 tracer=None, batch_rays=0):
        self.Source, self.RecMain, self.Room = sourceIn, receiverIn, roomIn
        self.COTime = cutoffTime / roomIn.Sound_speed(0)
        self._cutoff = float(cutoffTime)
        self._octaves = list(range(octaveRange[0], octaveRange[1] + 1))
        self.IS_Order = int(isOrderIn)
        self.Raycount = int(rayCountIn)
        self.precision = A.PREC_FP64 if precision == "fp64" else A.PREC_FP32
        self.mode = A.MODE_SPECULAR if mode == "specular" else A.MODE_STOCHASTIC
        self.seed = int(time.time_ns() & 0xFFFFFFFFFFFF) if seed is None else int(seed)
        self._tr = tracer or A.Tracer(device)
        self._own = tracer is None
        self._st = None
        self._lost_percent = 0.0
        self._stats = None
        self._started = False
        self._batch = int(batch_rays)

    # -- Simulation_Type
    def Sim_Type(self):
        return "Stochastic Ray Tracing"

    def Begin(self):
        r, s, rb = self.Room, self.Source, self.RecMain
        self._tr.upload_scene(r._tris, r._surf, r._alpha, r._scat, r._trans, r.grid)
        self._tr.set_source(s.Origin, s.SoundPower, s.dir_mode, s.dirs)
        if isinstance(rb, PachMapReceiver):
            # the medium and cut-off still come through the receiver set-up
            self._tr.set_receivers(rb.Origins[:1], rb.Radius, r.atten, r.c, r.rho * r.c, 1000, rb.SampleCT)
            self._tr.set_map(rb.Origins, rb.increment, rb.Directional, rb.SampleCT, rb.Time1Pt)
        else:
            if getattr(self._tr, "map_n", 0):
                self._tr.clear_map()
            self._tr.set_receivers(rb.Origins, rb.radius, r.atten, r.c, r.rho * r.c, rb.SampleRate,
                                   rb.SampleCT)
        self._st = time.time()
        self._started = True
        self._tr.trace(rays=self.Raycount, cutoff_m=self._cutoff, seed=self.seed, mode=self.mode,
                       precision=self.precision, oct_lo=self._octaves[0], oct_hi=self._octaves[-1],
                       is_order=self.IS_Order, batch_rays=self._batch, async_=True)

    def ThreadState(self):
        if not self._started:
            return Stopped
        return Running if self._tr.poll()["running"] else Stopped

    def ProgressMsg(self):
        p = self._tr.poll()
        if self.Raycount < 1:
            return (f"Revolution {-(-p['rays_done'] // 80)} - Ray {p['rays_done']} : sub {p['octave_rays']} - "
                    f"Convergence not yet reached...")
        if p["rays_done"] == 0 and p["running"]:
            return "Preparing Threads"
        return (f"Calculating Ray {p['rays_done']} of {p['rays_total']} "
                f"(sub-ray {p['octave_rays']}).")

    def Conclude_Simulation(self):
        """Cancels the trace; the rays already traced are kept (TaskAll:254-257, :276-280)."""
        self._tr.cancel()

    def Combine_ThreadLocal_Results(self):
        rc = self._tr.wait()
        if rc not in (0, -5):
            raise A.PachError(rc, A.lib().pach_last_error(self._tr.h).decode())
        self._stats = self._tr.stats()
        n = max(self._stats["rays_done"], 1)
        if isinstance(self.RecMain, PachMapReceiver):
            self.RecMain.planes = self._tr.map_planes(scaled=False)
        else:
            self.RecMain.planes = self._tr.planes(scaled=False)
        self.RecMain.Scale(n)
        self._lost_percent = self._tr.percent_lost()

    @property
    def GetReceiver(self):
        return self.RecMain

    @property
    def PercentLost(self):
        return self._lost_percent

    @property
    def Convergence_Report(self):
        """Convergence mode: receivers checked, revolutions and outcome (TaskAll:724-727)."""
        if self.Raycount >= 1 or self._stats is None:
            return None
        st = self._stats
        return {"visible_receiver": st["conv_receiver_visible"], "occluded_receiver": st["conv_receiver_occluded"],
                "revolutions": st["revolutions"], "converged": bool(st["converged"])}

    @property
    def Stats(self):
        return self._stats

    def close(self):
        if self._own:
            self._tr.close()


# ---------------------------------------------------------------- deterministic direct sound
class Direct_Sound(Simulation_Type):
    """Deterministic direct sound of a point source (Classes_Direct.cs:31-1010).

    Begin computes, for every receiver, whether the path from the source is
    open (Check_Validity, continuing through transmissive polygons), the
    arrival time d / c and the energy per band P 10^(-0.1 a d) transmod /
    (4 pi d^2) at sample 0 of a 100-sample direct-sound window
    (pach_direct_sound).  As in the reference the energy is computed for
    blocked paths too; IsOccluded (sic) returns True when the path is OPEN.
    Screen (edge-diffraction) attenuation of blocked paths is not provided."""

    def __init__(self, Src_in, Rec_in, Room_in, Octaves, ScreenCalc=False, SPL_Only=False):
        if ScreenCalc:
            raise NotImplementedError("screen attenuation of blocked direct paths is not provided")
        self.Src, self.Rec, self.Room = Src_in, Rec_in, Room_in
        self.Oct_choice = Octaves
        self.SPL_Only = bool(SPL_Only)
        self.Delay_ms = 0.0
        self.SampleFreq = Rec_in.SampleRate
        self._done = False

    def Sim_Type(self):
        return "Deterministic Direct Sound"

    def Begin(self):
        r = self.Room
        room = _RoomView(r._tris, r._surf, r._trans, self.Src.Origin, self.Rec.Origins)
        E, t, aud = A.direct_sound(room, self.Src.SoundPower, r.atten, r.c)
        n = 1 if self.SPL_Only else 100  # samples = ceil((max - min) fs / c) + 100 for one point
        self.Io = np.zeros((len(self.Rec.Origins), 8, n))
        self.Io[:, :, 0] = E
        self.Time_Pt = t + self.Delay_ms
        self.Validity = aud
        self._done = True

    def ThreadState(self):
        return Stopped

    def ProgressMsg(self):
        return "Calculating direct sound: complete" if self._done else "Calculating direct sound..."

    def Combine_ThreadLocal_Results(self):
        pass

    def IsOccluded(self, id):
        return bool(self.Validity[id])

    def Time(self, Rec_ID):
        return float(self.Time_Pt[Rec_ID])

    def Min_Time(self, Rec_ID):
        return float(self.Time_Pt[Rec_ID])

    def EnergyValue(self, octave, Rec_ID):
        return self.Io[Rec_ID, octave].copy()

    def EnergySum(self, Rec_ID, t):
        return float(self.Io[Rec_ID, :, t].sum()) if t < self.Io.shape[2] else 0.0


class Specular_Path:
    """One valid specular reflection path (Classes_ImageSource.cs:1891-2100).

    Path[0]: source, reflection points, receiver; TravelTime = length / c;
    Energy(Octave) = [PathEnergy[Octave]] (Octave 8: the band sum);
    prms = sqrt(PathEnergy rho c)."""

    def __init__(self, rec, order, seq, polys, pts, length, time_s, energy, rho_c, src_id=0, direct_time=0.0):
        self.Receiver = int(rec)
        self.Sequence = [int(x) for x in seq[:order]]
        self.Polygons = [int(x) for x in polys[:order]]
        self.ValidPath = np.array(pts[: order + 2], dtype=np.float64)
        self.Length = float(length)
        self.Time = float(time_s)
        self.PathEnergy = np.array(energy, dtype=np.float64)
        self.prms = np.sqrt(self.PathEnergy * rho_c)
        self.Identifier = self._identify(src_id, direct_time)

    def _identify(self, src_id, direct_time):
        # Identify (:1999-2034)
        return (f"S{src_id:03d}-Order {len(self.Sequence):03d}: {round((self.Time - direct_time) * 1000)} ms. "
                + "".join(f"{d} " for d in self.Sequence))

    def __str__(self):
        return self.Identifier

    @property
    def Path(self):
        return [self.ValidPath]

    @property
    def TravelTime(self):
        return self.Time

    @property
    def Reflection_Sequence(self):
        return list(self.Sequence)

    def Energy(self, Octave, SampleFrequency=44100):
        if Octave < 8:
            return np.array([self.PathEnergy[Octave]])
        return np.array([self.PathEnergy.sum()])

    def Set_Power(self, factor):
        self.PathEnergy = self.PathEnergy * np.asarray(factor, dtype=np.float64)


class ImageSourceData(Simulation_Type):
    """Image-source method (Classes_ImageSource.cs:30-1836) for planar objects
    without edge diffraction, validated on the GPU (pach_image_sources).

    Begin enumerates every sequence of 1..MaxOrder objects (repeats allowed,
    TraverseOrders :266-348) for every receiver centre; Paths[rec] holds the
    valid Specular_Paths in the reference's depth-first order.  Edge
    diffraction and curved objects (Compound_Path) are not provided."""

    def __init__(self, Source, Receiver, Direct, Rm, MaxOrder_in, SourceID_in=0, Edge_Diffraction=False,
                 *, device=0, tracer=None):
        if Edge_Diffraction:
            raise NotImplementedError("edge diffraction (Compound_Path) is not provided")
        self.Src, self.Rec, self.Direct, self.Room = Source, Receiver, Direct, Rm
        self.MaxOrder = int(MaxOrder_in)
        self.SrcNo = int(SourceID_in)
        self.Diffraction = False
        self._tr = tracer
        self._device = device
        self.ValidPaths = [[] for _ in range(len(Receiver.Origins))]
        self._done = False
        self.raw = None

    def Sim_Type(self):
        return "Image Source"

    def Begin(self):
        r, s, rb = self.Room, self.Src, self.Rec
        tr = self._tr or A.Tracer(self._device)
        try:
            tr.upload_scene(r._tris, r._surf, r._alpha, r._scat, r._trans, r.grid)
            tr.set_source(s.Origin, s.SoundPower, s.dir_mode, s.dirs)
            tr.set_receivers(rb.Origins, rb.radius, r.atten, r.c, r.rho * r.c, rb.SampleRate, max(rb.SampleCT, 1))
            P = tr.image_sources(self.MaxOrder, r._object, r.ObjectCount)
        finally:
            if self._tr is None:
                tr.close()
        self.raw = P
        rho_c = r.rho * r.c
        self.ValidPaths = [[] for _ in range(len(rb.Origins))]
        for p in P:
            dt = self.Direct.Min_Time(int(p["receiver"])) if self.Direct is not None else 0.0
            self.ValidPaths[int(p["receiver"])].append(
                Specular_Path(p["receiver"], p["order"], p["seq"], p["poly"], p["pts"], p["length"], p["time_s"],
                              p["energy"], rho_c, self.SrcNo, dt))
        self._done = True

    def ThreadState(self):
        return Stopped

    def ProgressMsg(self):
        return "Calculating Image Set: complete" if self._done else "Calculating Image Set..."

    def Combine_ThreadLocal_Results(self):
        pass

    @property
    def Paths(self):
        return self.ValidPaths

    def Count(self, rec_id):
        return len(self.ValidPaths[rec_id])

    def Set_Power(self, Factor):
        for paths in self.ValidPaths:
            for p in paths:
                p.Set_Power(Factor)


class _RoomView:
    """The fields of rooms.Room that pach_direct_sound reads."""

    def __init__(self, verts, surf, trans, source, receivers):
        self.verts, self.surf, self.trans, self.source, self.receivers = verts, surf, trans, source, receivers


class Pach_SP:
    """Pach_SP (Pach_SP.cs) signal processing used by the auralisation path."""

    @staticmethod
    def ETCToFilter(Octave_PRMS, SWL, sample_frequency=44100, sample_frequency_out=44100, device=0):
        """Pach_SP.ETCToFilter (Pach_SP.cs:1584-1676) on the GPU: the sum over the
        octave bands (and a 16 kHz band from the 8 kHz one) of each band's
        |envelope| 10^((120 - SWL) / 20) convolved with its band-limited
        carrier; length T + 16383.  sample_frequency_out is unused, as in the
        reference."""
        with A.Tracer(device) as tr:
            return tr.etc_to_filter(np.asarray(Octave_PRMS, dtype=np.float64), SWL, sample_frequency)


class IR_Construction:
    """Classes_Math.cs, IR_Construction."""

    @staticmethod
    def ETCurve(Direct, ISData, RTData, CO_Time_ms, Sampling_Frequency, Octave, Rec_ID, Src_ID=0,
                Start_at_Zero=False):
        """Energy-time curve of one receiver and band (Octave 8: all bands) from
        the direct sound, image sources and the ray tracer's scaled histograms
        (Classes_Math.cs:1417-1474).  Each image-source path adds its energy at
        ceil(fs (TravelTime + delay)) when that bin lies inside (0, n - 1).
        Where the reference would index past the end of the histogram, the
        direct sound's bins are dropped."""
        Direct = list(Direct) if Direct is not None else [None]
        ISData = list(ISData) if ISData is not None else [None]
        delay = Direct[Src_ID].Delay_ms if Direct[0] is not None else 0.0
        if RTData[Src_ID] is not None:
            H = RTData[Src_ID].GetEnergyHistogram(Octave, delay, Rec_ID).copy()
        else:
            H = np.zeros(int((CO_Time_ms + int(delay)) * 0.001 * Sampling_Frequency))
        D = Direct[Src_ID]
        if D is not None and D.IsOccluded(Rec_ID):
            start = 0 if Start_at_Zero else int(np.ceil((delay * 0.001 + D.Time(Rec_ID)) * Sampling_Frequency))
            n = D.Io.shape[2]
            for i in range(n):
                v = D.EnergySum(Rec_ID, i) if Octave == 8 else D.EnergyValue(Octave, Rec_ID)[i]
                if start + i < H.size:
                    H[start + i] += v
        if ISData[Src_ID] is not None:
            for value in ISData[Src_ID].Paths[Rec_ID]:
                place = int(np.ceil(Sampling_Frequency * (value.TravelTime + delay * 0.001)))
                if 0 < place < H.size - 1:
                    e = value.Energy(Octave, Sampling_Frequency)
                    for t in range(e.size):
                        if place + t < H.size - 1:
                            H[place + t] += e[t]
        return H


# ---------------------------------------------------------------- decay-curve integration
class AcousticalMath:
    """Classes_Math.cs:272-585, FP64 in libpach.so."""

    @staticmethod
    def SoundSpeed(Air_Temp):
        return 331 + 0.6 * Air_Temp

    @staticmethod
    def Schroeder_Integral(etc):
        return A.schroeder_integral(etc)

    @staticmethod
    def T_X(Schroeder_Integral, Decay_Depth, sample_f):
        return A.t_x(Schroeder_Integral, Decay_Depth, sample_f)

    @staticmethod
    def EarlyDecayTime(Schroeder_Integral, sample_f):
        return A.early_decay_time(Schroeder_Integral, sample_f)

    @staticmethod
    def SPL_Intensity(Intensity):
        """Classes_Math.cs:97-102: 10 log10(I / 1e-12), 0 below 0 dB."""
        with np.errstate(divide="ignore"):
            spl = 10 * np.log10(Intensity / 1e-12)
        return 0.0 if not spl >= 0 else float(spl)

    @staticmethod
    def Clarity(etc, sample_f, seconds, startTime, pressure=False):
        return A.clarity(etc, sample_f, seconds, startTime, pressure)
