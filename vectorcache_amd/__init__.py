"""vectorcache_amd: MI355X (gfx950) TurboQuant encode + flat search, and the
landmark index built on it.

The GPU library is loaded lazily by ``vectorcache_amd.index`` and
``vectorcache_amd.landmark`` (each raises ImportError if it is not built):
``from vectorcache_amd.index import TurboQuantIndex`` (the plain flat index),
``from vectorcache_amd.landmark import LandmarkIndex`` (the reference's
LandmarkIndex, which its TurboQuantIndex aliases). ``vectorcache_amd.formats``
reads and writes the NPY and VCEX files. ``vectorcache_amd.synthetic``
(seeded workloads) and ``vectorcache_amd.build`` need no GPU.
"""

__all__ = ["index", "landmark", "formats", "synthetic", "build"]
