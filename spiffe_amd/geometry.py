"""Host-side geometry over the C ABI: boundary files, the boundary drawing and
the metal flags (spf_geometry_parse / spf_geometry_draw / spf_geometry_flags,
the library's restatement of src/geometry.c:21-369 and src/grid_drawing.c).
Arrays use the reference layout [iz][ir], shape (nz+1, nr+1).
"""
import ctypes

import numpy as np

from . import abi
from .abi import FL_IS_METAL, FL_IS_SURFACE, FL_LAST_ACTIVE_PT

DBL_MAX = np.finfo(np.float64).max


def _err():
    return ctypes.create_string_buffer(512)


def parse_boundary(text, zr_factor=1.0):
    """The &po/&point namelists of a boundary file -> ctypes array of spf_boundary_point."""
    lib = abi.load()
    data = text.encode() if isinstance(text, str) else bytes(text)
    n = ctypes.c_int64()
    e = _err()
    rc = lib.spf_geometry_parse(data, float(zr_factor), None, 0, ctypes.byref(n), e, len(e))
    if rc:
        raise ValueError(e.value.decode())
    pts = (abi.BoundaryPoint * n.value)()
    rc = lib.spf_geometry_parse(data, float(zr_factor), pts, n.value, ctypes.byref(n), e, len(e))
    if rc:
        raise ValueError(e.value.decode())
    return pts


def draw(nz, nr, zmin, zmax, rmax, points):
    """Flags (with the field bits), boundary potentials (PsiImposed) and material ids."""
    lib = abi.load()
    flags = np.zeros((nz + 1, nr + 1), dtype=np.int16)
    psi = np.zeros((nz + 1, nr + 1), dtype=np.float64)
    mat = np.zeros((nz + 1, nr + 1), dtype=np.int16)
    e = _err()
    rc = lib.spf_geometry_draw(int(nz), int(nr), float(zmin), float(zmax), float(rmax), points, len(points),
                               flags.ctypes.data_as(ctypes.POINTER(ctypes.c_int16)),
                               psi.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                               mat.ctypes.data_as(ctypes.POINTER(ctypes.c_int16)), e, len(e))
    if rc:
        raise ValueError(e.value.decode())
    return flags, psi, mat


def boundary_file(path, nz, nr, zmin, zmax, rmax, zr_factor=1.0):
    """define_geometry's boundary file drawn on the grid: (flags, PsiImposed, materialID)."""
    with open(path) as fp:
        return draw(nz, nr, zmin, zmax, rmax, parse_boundary(fp.read(), zr_factor))


def polygon_points(poly, potentials=None):
    """A closed polyline of (z, r) vertices as nt=1 boundary points."""
    pts = (abi.BoundaryPoint * len(poly))()
    for k, (z, r) in enumerate(poly):
        pts[k].nt = 1
        pts[k].x, pts[k].y = z, r
        pts[k].theta = DBL_MAX
        pts[k].potential = potentials[k] if potentials else 0.0
    return pts


def derive_flags(marks):
    """Adds the E/B zero bits and FL_LAST_ACTIVE_PT (src/geometry.c:317-362) to a
    copy of an array holding only FL_IS_SURFACE / FL_IS_INSIDE marks."""
    f = np.array(marks, dtype=np.int16, copy=True, order="C")
    rc = abi.load().spf_geometry_flags(f.shape[0] - 1, f.shape[1] - 1, f.ctypes.data_as(ctypes.POINTER(ctypes.c_int16)))
    if rc:
        raise ValueError("bad flag array shape %r" % (f.shape,))
    return f


def pillbox_marks(nz, nr):
    """Metal marks of a closed cylinder whose walls are the grid edges z=zmin
    (iz=0), z=zmax (iz=nz-1) and r=rmax (ir=nr-1), set directly (a drawn
    boundary on the grid edges would be trimmed away, src/grid_drawing.c:1165-1192)."""
    f = np.zeros((nz + 1, nr + 1), dtype=np.int16)
    f[0, :nr] = FL_IS_SURFACE
    f[nz - 1, :nr] = FL_IS_SURFACE
    f[:nz, nr - 1] = FL_IS_SURFACE
    return f


def polygon_marks(nz, nr, zmin, zmax, rmax, poly, potentials=None):
    """Surface/inside marks of a closed polyline drawn, filled and trimmed as the
    reference draws a boundary file of nt=1 points."""
    flags, _, _ = draw(nz, nr, zmin, zmax, rmax, polygon_points(poly, potentials))
    return (flags & FL_IS_METAL).astype(np.int16)


def last_active_cuts(flags):
    """Per-row loop cut-offs the device uses: first iz >= start flagged
    FL_LAST_ACTIVE_PT, else nz-1 (src/integrate.c:345-346 and siblings)."""
    nz = flags.shape[0] - 1
    out = {}
    for start in (0, 1):
        cut = np.full(flags.shape[1], nz - 1, dtype=np.int64)
        for ir in range(flags.shape[1]):
            hit = np.nonzero(flags[start:nz, ir] & FL_LAST_ACTIVE_PT)[0]
            if hit.size:
                cut[ir] = hit[0] + start
        out[start] = cut
    return out
