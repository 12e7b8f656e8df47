"""Persistent cache of kernels compiled at run time (``src/jht/_cache.py:25-40``).

jht compiles its transforms with XLA at first use and offers a persistent on-disk cache
for those compilations.  jht_amd's own kernels are compiled ahead of time for gfx950 into
``libjhtx.so``, so there is nothing of its own to cache; the only kernels built at run
time are rocFFT's (the plain 2-D FFTs of the off-grid pair, through ``torch.fft``), which
rocFFT compiles on first use and can keep in an on-disk cache.
"""

from __future__ import annotations

import os


def enable_compilation_cache(directory: str, *, min_compile_time_secs: float = 1.0) -> None:
    """Keep runtime-compiled kernels in ``directory`` across processes.

    Points rocFFT's kernel cache (``ROCFFT_RTC_CACHE_PATH``) at a file in ``directory``
    (created if missing).  Call once at program start, before the first off-grid
    transform, as with jht.  ``min_compile_time_secs`` is accepted for compatibility and
    has no effect: rocFFT caches every kernel it compiles.  Numerics are unaffected.
    """
    if float(min_compile_time_secs) < 0:
        raise ValueError("min_compile_time_secs must be >= 0")
    d = os.path.expanduser(directory)
    os.makedirs(d, exist_ok=True)
    os.environ["ROCFFT_RTC_CACHE_PATH"] = os.path.join(d, "rocfft_kernel_cache.db")


__all__ = ["enable_compilation_cache"]
