"""Uniform streams that drive FRI systematic sampling.

The reference draws exactly one uniform per systematic compression as
``rand()/RAND_MAX`` from the C library's global generator
(``lin_space_add_const``, src/fri_utils.cpp:45-58).  ``GlibcUniforms`` calls
that same C function (it does not reimplement it), so a solve seeded with
``srand(seed)`` consumes the same stream as the reference in a process that
called ``srand(seed)``.  Note glibc: srand(0), srand(1) and no seeding give
the same stream.

``PhiloxUniforms`` (counter-based, stateless per (seed, draw)) serves
independent replica trajectories, where glibc's single global state and
colliding seeds are unsuitable.
"""
import ctypes

import numpy as np

RAND_MAX = 2147483647

_libc = None


def libc():
    global _libc
    if _libc is None:
        _libc = ctypes.CDLL(None)
        _libc.rand.restype = ctypes.c_int
        _libc.srand.argtypes = [ctypes.c_uint]
    return _libc


def draw_libc():
    """One uniform from the C library's global generator, rand()/RAND_MAX in [0,1]."""
    return libc().rand() / float(RAND_MAX)


class GlibcUniforms:
    """Seeds the C library generator with ``srand(seed)`` and draws from it."""

    def __init__(self, seed=None):
        if seed is not None:
            libc().srand(int(seed))
        self.count = 0

    def __iter__(self):
        return self

    def __next__(self):
        self.count += 1
        return draw_libc()

    def state(self):
        return {"kind": "glibc", "count": self.count}


class PhiloxUniforms:
    """numpy Philox4x32 stream: independent per seed, resumable by draw count."""

    def __init__(self, seed, count=0):
        self.seed = int(seed)
        self.count = int(count)
        self._bg = np.random.Philox(key=self.seed)
        if count:
            self._bg.advance(count)
        self._gen = np.random.Generator(self._bg)

    def __iter__(self):
        return self

    def __next__(self):
        self.count += 1
        return float(self._gen.random())

    def state(self):
        return {"kind": "philox", "seed": self.seed, "count": self.count}


class FixedUniforms:
    """A given sequence of uniforms (tests, replays)."""

    def __init__(self, values):
        self._it = iter(list(values))
        self.count = 0

    def __iter__(self):
        return self

    def __next__(self):
        self.count += 1
        return float(next(self._it))
