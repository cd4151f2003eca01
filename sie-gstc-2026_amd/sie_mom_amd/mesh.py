"""Mesh handle: host arrays plus the device-resident sie_mesh_t."""
import ctypes

import numpy as np

from . import _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class Mesh:
    """A closed, outward-oriented triangle mesh (MeshData,
    include/sie_mom/mesh/spherical_surface.hpp:10-15) uploaded to the GPU."""

    def __init__(self, vertices, faces, edges, face_edges):
        self.vertices = np.ascontiguousarray(vertices, np.float64)
        self.faces = np.ascontiguousarray(faces, np.int32)
        self.edges = np.ascontiguousarray(edges, np.int32)
        self.face_edges = np.ascontiguousarray(face_edges, np.int32)
        self._h = None

    @property
    def nf(self):
        return len(self.faces)

    @property
    def ne(self):
        return len(self.edges)

    @property
    def handle(self):
        if self._h is None:
            h = ctypes.c_void_p()
            _lib.check(_lib.lib.sie_mesh_create(
                _p(self.vertices), len(self.vertices), _p(self.faces), len(self.faces),
                _p(self.face_edges), len(self.edges), ctypes.byref(h)))
            self._h = h
        return self._h

    def __del__(self):
        if getattr(self, "_h", None) is not None:
            try:
                _lib.lib.sie_mesh_destroy(self._h)
            except Exception:
                pass

    @staticmethod
    def _generate(fn, *args):
        counts = np.zeros(3, np.int32)
        _lib.check(fn(*args, None, None, None, None, _p(counts)))
        nv, nf, ne = (int(c) for c in counts)
        V = np.zeros((nv, 3))
        F = np.zeros((nf, 3), np.int32)
        E = np.zeros((ne, 2), np.int32)
        FE = np.zeros((nf, 3), np.int32)
        _lib.check(fn(*args, _p(V), _p(F), _p(E), _p(FE), _p(counts)))
        return Mesh(V, F, E, FE)

    @staticmethod
    def icosphere(radius, level):
        """SphericalSurface(radius, level): 20*4^level faces."""
        f = _lib.lib.sie_icosphere
        f.argtypes = [ctypes.c_float, ctypes.c_int] + [ctypes.c_void_p] * 5
        f.restype = ctypes.c_int
        return Mesh._generate(f, float(radius), int(level))

    @staticmethod
    def spindle(chord, arc_radius, n_rings):
        """Revolution of a circular arc (radius arc_radius) about its chord."""
        f = _lib.lib.sie_spindle
        f.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_int] + [ctypes.c_void_p] * 5
        f.restype = ctypes.c_int
        return Mesh._generate(f, float(chord), float(arc_radius), int(n_rings))

    @staticmethod
    def from_obj(path):
        """The strict OBJ loader (ObjMeshLoader, obj_mesh_loader.cpp:83-386):
        vertex and face order preserved, closed-surface checks, RWG edges in
        first-encounter order; raises SieError with the loader's message."""
        f = _lib.lib.sie_obj_load
        f.argtypes = [ctypes.c_char_p] + [ctypes.c_void_p] * 5
        f.restype = ctypes.c_int
        return Mesh._generate(f, str(path).encode())
