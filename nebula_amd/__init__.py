"""MI355X-native electron tracking through triangle-meshed samples: a drop-in
for Nebula's GPU tracking path (drivers/gpu of the reference) built on
hand-written HIP kernels for gfx950 behind a C ABI (include/nebula.h).

Submodules: `api` (Device, Geometry, Material, Driver over libnebula.so),
`workloads` (synthetic geometries, beams and material tables; .tri/.pri
formats).
"""
__version__ = "0.1.0"
