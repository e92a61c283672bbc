"""vectorcache_amd: MI355X (gfx950) TurboQuant encode + flat search.

The GPU library is loaded lazily by ``vectorcache_amd.index`` (and raises
ImportError if it is not built): ``from vectorcache_amd.index import
TurboQuantIndex``. ``vectorcache_amd.synthetic`` (seeded workloads) and
``vectorcache_amd.build`` need no GPU.
"""

__all__ = ["index", "synthetic", "build"]
