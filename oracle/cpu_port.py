"""ctypes binding of oracle/cpu/zbcpu.cpp (built into oracle/_build/libzbcpu.so): the indexing path
as a float32 C++17 / OpenMP program on the host cores, the CPU restatement of the reference
engine that bench.py times beside the GPU.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py): the tables it takes (Gauss-Legendre rings,
Legendre and Wigner tables, the master's weights) are set-up inputs handed in by the caller.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "libzbcpu.so")
f32p = ctypes.POINTER(ctypes.c_float)
i32 = ctypes.c_int32


class Band(ctypes.Structure):
    _fields_ = [("L", i32), ("n", i32), ("ring", f32p), ("wplm", f32p), ("dq", f32p)]


class Phase(ctypes.Structure):
    _fields_ = [("fold", i32), ("nc", i32), ("mu_max", i32), ("l_step", i32), ("w", f32p)]


class Frame(ctypes.Structure):
    _fields_ = [("c0", ctypes.c_float * 3), ("c1", ctypes.c_float * 3), ("c2", ctypes.c_float * 3), ("h", i32), ("w", i32),
                ("mask", ctypes.c_float * 3)]


class Master(ctypes.Structure):
    _fields_ = [("side", i32), ("sheets", f32p), ("grad", f32p)]


class Opts(ctypes.Structure):
    _fields_ = [("iters", i32), ("max_step", ctypes.c_float), ("scale0", ctypes.c_float), ("scale_min", ctypes.c_float),
                ("scale_max", ctypes.c_float), ("nmi_bins", i32)]


def build():
    """Compile the library with its Makefile (g++, OpenMP); raises on a compiler error."""
    r = subprocess.run(["make", "-s", "-C", os.path.join(HERE, "cpu")], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"building {LIB} failed:\n{r.stderr}")
    return LIB


_lib = None


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError(f"{LIB} is missing: build it with oracle.cpu_port.build() (or __graft_entry__.build())")
        lib = ctypes.CDLL(LIB)
        lib.zbc_threads.restype = i32
        lib.zbc_rden.argtypes = [ctypes.POINTER(Band), ctypes.POINTER(Phase), ctypes.POINTER(Phase), ctypes.POINTER(Frame), f32p, f32p]
        lib.zbc_rden.restype = i32
        lib.zbc_index.argtypes = [ctypes.POINTER(Band), ctypes.POINTER(Phase), ctypes.POINTER(Frame), ctypes.POINTER(Master),
                                  ctypes.POINTER(Opts), f32p, f32p, i32, f32p, f32p, f32p, i32]
        lib.zbc_index.restype = i32
        _lib = lib
    return _lib


def _p(a):
    return a.ctypes.data_as(f32p)


class Port:
    """One phase at one bandwidth and detector frame. band: an object with L, n, ring [n][2],
    wplm [n/2][pairs], dq [L][L][L]; phase: fold, nc, mu_max, l_step, w_f, w_f2 ([mc][L][L][2]);
    fr: the oracle's frame dict (c0, c1, c2, h, w, mask); sheets [2][side][side] and grad
    [2][side][side][4] float32."""

    def __init__(self, band, phase, fr, sheets, grad, side, iters=5, max_step=np.radians(2.0), scale0=1.0, scale_min=0.02,
                 scale_max=1e4, nmi_bins=16):
        self.lib = load()
        self._keep = [np.ascontiguousarray(x, np.float32) for x in (band.ring, band.wplm, band.dq, phase.w_f, phase.w_f2, sheets, grad)]
        ring, wplm, dq, wf, wf2, sh, gr = self._keep
        self.band = Band(band.L, band.n, _p(ring), _p(wplm), _p(dq))
        self.phase = Phase(phase.fold, phase.nc, phase.mu_max, phase.l_step, _p(wf))
        self.phase2 = Phase(phase.fold, phase.nc, phase.mu_max, phase.l_step, _p(wf2))
        self.frame = Frame()
        self.frame.c0[:] = [float(x) for x in fr["c0"]]
        self.frame.c1[:] = [float(x) for x in fr["c1"]]
        self.frame.c2[:] = [float(x) for x in fr["c2"]]
        self.frame.h, self.frame.w = int(fr["h"]), int(fr["w"])
        self.frame.mask[:] = [float(x) for x in fr["mask"]]
        self.master = Master(int(side), _p(sh), _p(gr))
        self.opts = Opts(int(iters), float(max_step), float(scale0), float(scale_min), float(scale_max), int(nmi_bins))
        self.n, self.nc = band.n, phase.nc

    def rden(self, t):
        """The window's reciprocal denominator volume [n/2 + 1][nc][n] at detector position t."""
        out = np.empty((self.n // 2 + 1, self.nc, self.n), np.float32)
        t = np.ascontiguousarray(t, np.float32)
        if self.lib.zbc_rden(self.band, self.phase, self.phase2, self.frame, _p(t), _p(out)) != 0:
            raise RuntimeError("zbc_rden failed")
        return out

    def index(self, pats, items, rden, threads=0):
        """records [n][8] (q, NCC, NMI, step, 0) and peaks [n][8] (q, score, ...) of raw float32
        patterns [*][h][w] at items [n][4] (t, pattern index)."""
        pats = np.ascontiguousarray(pats, np.float32)
        items = np.ascontiguousarray(items, np.float32)
        rden = np.ascontiguousarray(rden, np.float32)
        n = len(items)
        rec = np.empty((n, 8), np.float32)
        pk = np.empty((n, 8), np.float32)
        if self.lib.zbc_index(self.band, self.phase, self.frame, self.master, self.opts, _p(pats), _p(items), n, _p(rden), _p(rec), _p(pk),
                              int(threads)) != 0:
            raise RuntimeError("zbc_index failed")
        return rec, pk


def threads():
    return int(load().zbc_threads())
