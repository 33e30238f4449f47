"""Per-kernel durations measured with HIP events inside libprt.

    with KernelTimes() as kt:
        ... calls into libprt ...
    kt.stats  # {kernel name: (launches, total ms)}
"""
from __future__ import annotations

import ctypes

import numpy as np

from ._lib import lib


class KernelTimes:
    def __init__(self):
        self.stats: dict[str, tuple[int, float]] = {}

    def __enter__(self):
        lib().prt_profile_read(None, 0, None, None, 0)  # drop stale records
        lib().prt_profile_enable(1)
        return self

    def __exit__(self, *exc):
        lib().prt_profile_enable(0)
        names = ctypes.create_string_buffer(1 << 16)
        cnt = np.zeros(256, np.int64)
        ms = np.zeros(256, np.float64)
        k = lib().prt_profile_read(names, len(names), ctypes.c_void_p(cnt.ctypes.data),
                                   ctypes.c_void_p(ms.ctypes.data), 256)
        if k < 0:
            raise RuntimeError(f"prt_profile_read failed ({k})")
        labels = names.raw.split(b"\0")[:k]
        for i, lab in enumerate(labels):
            self.stats[lab.decode()] = (int(cnt[i]), float(ms[i]))
        return False

    def avg_ms(self, name: str) -> float:
        c, t = self.stats[name]
        return t / c

    def dominant(self) -> str:
        return max(self.stats, key=lambda k: self.stats[k][1])
