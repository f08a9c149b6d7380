"""MI355X-native stochastic split ray tracer with Pachyderm's Complete_Trace API.

    from pachydermacoustic_universal_amd import SplitRayTracer, Polygon_Scene, ...

The compute path is libpach.so (gfx950 HIP kernels behind the C ABI of
include/pach.h); importing the tracer classes does not require a GPU, but
running a trace does, and fails loudly without one.
"""
from . import rooms  # noqa: F401
from .pachyderm import (  # noqa: F401
    AcousticalMath,
    Basic_Material,
    Direct_Sound,
    GeodesicSource,
    IR_Construction,
    ImageSourceData,
    Lambert_Scattering,
    Pach_SP,
    PachMapReceiver,
    Polygon_Scene,
    RandomSource,
    Receiver_Bank,
    Running,
    Simulation_Type,
    Specular_Path,
    SplitRayTracer,
    Stopped,
    SuppliedDirectionSource,
)

__all__ = ["AcousticalMath", "Basic_Material", "Direct_Sound", "GeodesicSource", "IR_Construction",
           "ImageSourceData", "Lambert_Scattering", "Pach_SP", "PachMapReceiver", "Specular_Path",
           "Polygon_Scene", "RandomSource", "Receiver_Bank", "Running", "Simulation_Type",
           "SplitRayTracer", "Stopped", "SuppliedDirectionSource", "rooms"]
