"""match-series on MI355X: Python-side helpers of the gfx950 registration library.

The product is libmsgpu.so (C ABI in include/ms_gpu.h) and the matchSeries drop-in CLI; this package
holds the build script, the ctypes binding used by tests and benchmarks, and the synthetic series
generator.
"""
