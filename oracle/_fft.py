"""FFT backend of the oracle: scipy.fft on every host core available to this process (identical
results to numpy.fft's definition; unnormalised forward, 1/n inverse)."""
import os

import scipy.fft as _sf

WORKERS = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)


def set_workers(n: int):
    global WORKERS
    WORKERS = max(1, int(n))


def fft(a, axis=-1):
    return _sf.fft(a, axis=axis, workers=WORKERS)


def ifft(a, axis=-1):
    return _sf.ifft(a, axis=axis, workers=WORKERS)


def fftn(a, s=None, axes=None):
    return _sf.fftn(a, s=s, axes=axes, workers=WORKERS)


def ifftn(a, axes=None):
    return _sf.ifftn(a, axes=axes, workers=WORKERS)
