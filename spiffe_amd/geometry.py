"""Host-side geometry: metal masks and the derived field flags.

derive_flags() restates the flag derivation of src/geometry.c:317-362 on a
numpy array in the reference layout ([iz][ir], shape (nz+1, nr+1)).  The
rasterizers here mark which grid points are metal for simple shapes (the
rectangle of a closed pillbox, or a polygon given in (z, r)); the reference's
own drawing routines (src/grid_drawing.c) are a separate, larger piece.
"""
import numpy as np

from .abi import (FL_BPHI_ZERO, FL_BR_ZERO, FL_BZ_ZERO, FL_EPHI_ZERO, FL_ER_ZERO, FL_EZ_ZERO, FL_IS_INSIDE,
                  FL_IS_METAL, FL_IS_SURFACE, FL_LAST_ACTIVE_PT)


def derive_flags(flags):
    """Adds the E/B zero bits and FL_LAST_ACTIVE_PT to an array holding only
    FL_IS_SURFACE / FL_IS_INSIDE marks.  src/geometry.c:319-362."""
    f = np.array(flags, dtype=np.int16, copy=True)
    nzp1, nrp1 = f.shape
    nz, nr = nzp1 - 1, nrp1 - 1
    metal = (f & FL_IS_METAL) != 0
    up = np.zeros_like(metal)          # metal at iz-1
    up[1:, :] = metal[:-1, :]
    left = np.zeros_like(metal)        # metal at ir-1
    left[:, 1:] = metal[:, :-1]
    diag = np.zeros_like(metal)        # metal at (iz-1, ir-1)
    diag[1:, 1:] = metal[:-1, :-1]
    ir0 = np.zeros_like(metal)
    ir0[:, 0] = True
    add = np.zeros(f.shape, dtype=np.int32)
    add |= np.where(metal, FL_EPHI_ZERO, 0)
    add |= np.where(metal & up, FL_EZ_ZERO | FL_BR_ZERO, 0)
    add |= np.where(metal & (ir0 | left), FL_ER_ZERO | FL_BZ_ZERO, 0)
    add |= np.where(metal & up & (ir0 | (left & diag)), FL_BPHI_ZERO, 0)
    f = (f.astype(np.int32) | add).astype(np.int16)
    for ir in range(nr):
        iz = nz - 1
        while iz >= 0 and f[iz, ir] & FL_IS_INSIDE:
            iz -= 1
        if iz != 0:
            iz += 1
            if iz < nz - 1:
                f[iz, ir] |= FL_LAST_ACTIVE_PT
    return f


def pillbox_marks(nz, nr):
    """Metal marks of a closed cylinder whose walls are the grid edges
    z=zmin (iz=0), z=zmax (iz=nz-1) and r=rmax (ir=nr-1)."""
    f = np.zeros((nz + 1, nr + 1), dtype=np.int16)
    f[0, :nr] = FL_IS_SURFACE
    f[nz - 1, :nr] = FL_IS_SURFACE
    f[:nz, nr - 1] = FL_IS_SURFACE
    return f


class Drawing:
    """The reference's boundary drawing on the (nz x nr) lattice of grid points,
    src/grid_drawing.c: draw_line (:215-295) marks points along each segment by
    rounding to the nearest grid point (add_point, :445-485); fill() flood-fills
    vacuum from the axis (:889-1038), turns unreached points to FL_IS_INSIDE and
    inside points touching vacuum to FL_IS_SURFACE (:1040-1083); trim() clears
    single-width metal along the left, right and bottom edges (:1165-1192)."""

    def __init__(self, nz, nr, zmin, zmax, rmax):
        self.nx, self.ny = nz, nr
        self.xmin, self.dx = zmin, (zmax - zmin) / (nz - 1)
        self.ymin, self.dy = 0.0, rmax / (nr - 1)
        self.lattice = np.zeros((nz, nr), dtype=np.int16)
        self.potential = np.zeros((nz, nr), dtype=np.float64)
        self.n_outside = 0

    def add_point(self, x, y, pot):
        i = int((x - self.xmin) / self.dx + 0.5)
        j = int((y - self.ymin) / self.dy + 0.5)
        if 0 <= i < self.nx and 0 <= j < self.ny:
            self.lattice[i, j] = FL_IS_SURFACE
            self.potential[i, j] = pot
        else:
            self.n_outside += 1

    def draw_line(self, xi, yi, xf, yf, potential=0.0, start_potential=None):
        sp = potential if start_potential is None else start_potential
        dx, dy = self.dx, self.dy
        if abs(xi - xf) >= dx:
            slope = (yf - yi) / (xf - xi)
            if abs(yf - yi) < dy / 1e6:
                slope = 0.0
            intercept = yi - slope * xi
            _dx = dx
            if abs(slope * _dx) > abs(dy):
                _dx = dy * .5 / abs(slope)
            pslope = (potential - sp) / (xf - xi)
            _dx = 0.5 * _dx
            x = xi
            while True:
                self.add_point(x, slope * x + intercept, sp + pslope * (x - xi))
                x = x + _dx if xi <= xf else x - _dx
                if (xi <= xf and not x <= xf) or (xi > xf and not x >= xf):
                    break
            self.add_point(xf, yf, potential)
        elif abs(yi - yf) >= dy:
            slope = (xf - xi) / (yf - yi)
            if abs(xf - xi) < dx / 1e6:
                slope = 0.0
            intercept = xi - slope * yi
            pslope = (potential - sp) / (yf - yi)
            _dy = dy * 0.5
            y = yi
            while True:
                self.add_point(slope * y + intercept, y, sp + pslope * (y - yi))
                y = y + _dy if yi <= yf else y - _dy
                if (yi <= yf and not y <= yf) or (yi > yf and not y >= yf):
                    break
            self.add_point(xf, yf, potential)
        else:
            self.add_point(xf, yf, potential)

    def fill(self):
        from scipy import ndimage
        L = self.lattice
        nx, ny = self.nx, self.ny
        seed = None
        for i in range(nx):
            if L[i, 0] and not L[i, 1]:
                break
        else:
            raise ValueError("can't determine what points are in vacuum")
        if i > 0:
            seed = (i, 1)
        vac = np.zeros_like(L, dtype=bool)
        if seed is not None:
            interior = np.zeros_like(vac)
            interior[1:nx - 1, 1:ny - 1] = L[1:nx - 1, 1:ny - 1] == 0
            interior[seed] = True
            lab, _ = ndimage.label(interior)
            comp = lab == lab[seed]
            vac |= comp
            grow = np.zeros_like(vac)
            grow[1:, :] |= comp[:-1, :]
            grow[:-1, :] |= comp[1:, :]
            grow[:, 1:] |= comp[:, :-1]
            grow[:, :-1] |= comp[:, 1:]
            vac |= grow & (L == 0)
        L[:] = np.where(vac, 0, np.where(L == 0, FL_IS_INSIDE, L))
        # inside points that touch vacuum become surface, in the reference's scan order
        zero = L == 0
        nbz = np.zeros(L.shape, dtype=np.int32)
        for di in (-1, 0, 1):
            for dj in (-1, 0, 1):
                if di == 0 and dj == 0:
                    continue
                sh = np.zeros_like(zero)
                sh[max(0, -di):nx - max(0, di), max(0, -dj):ny - max(0, dj)] = \
                    zero[max(0, di):nx - max(0, -di), max(0, dj):ny - max(0, -dj)]
                nbz += sh
        cand = set(map(tuple, np.argwhere((L == FL_IS_INSIDE) & (nbz > 0))))
        order = sorted(c for c in cand if 1 <= c[0] < nx - 1 and 1 <= c[1] < ny - 1)
        import heapq
        heap = list(order)
        heapq.heapify(heap)
        seen = set(heap)
        while heap:
            i, j = heapq.heappop(heap)
            if L[i, j] != FL_IS_INSIDE:
                continue
            change, psum, nsum = 0, 0.0, 0
            for i1 in (-1, 0, 1):
                for j1 in (-1, 0, 1):
                    if i1 == 0 and j1 == 0:
                        continue
                    v = L[i + i1, j + j1]
                    if not v:
                        change += 1
                    elif v == FL_IS_SURFACE:
                        psum += self.potential[i + i1, j + j1]
                        nsum += 1
            if change == 8:
                L[i, j] = 0
                for i1 in (-1, 0, 1):
                    for j1 in (-1, 0, 1):
                        q = (i + i1, j + j1)
                        if q > (i, j) and 1 <= q[0] < nx - 1 and 1 <= q[1] < ny - 1 and q not in seen \
                                and L[q] == FL_IS_INSIDE:
                            seen.add(q)
                            heapq.heappush(heap, q)
            elif change:
                L[i, j] = FL_IS_SURFACE
                self.potential[i, j] = psum / nsum if nsum else 0.0

    def trim(self):
        L = self.lattice
        for j in range(self.ny):
            if L[0, j] and not L[1, j]:
                L[0, j] = 0
        i = self.nx - 1
        for j in range(self.ny):
            if L[i, j] and not L[i - 1, j]:
                L[i, j] = 0
        for i in range(self.nx):
            if L[i, 0] and not L[i, 1]:
                L[i, 0] = 0

    def marks(self):
        """(nz+1, nr+1) int16 array holding FL_IS_SURFACE/FL_IS_INSIDE marks."""
        f = np.zeros((self.nx + 1, self.ny + 1), dtype=np.int16)
        f[:self.nx, :self.ny] = self.lattice
        return f

    def potentials(self):
        p = np.zeros((self.nx + 1, self.ny + 1), dtype=np.float64)
        p[:self.nx, :self.ny] = self.potential
        return p


def polygon_marks(nz, nr, zmin, zmax, rmax, poly, potentials=None):
    """Metal marks for a closed polyline of (z, r) vertices drawn, filled and
    trimmed as the reference draws a boundary of nt=1 line segments."""
    d = Drawing(nz, nr, zmin, zmax, rmax)
    pot = potentials or [0.0] * len(poly)
    d.add_point(poly[0][0], poly[0][1], pot[0])
    for k in range(1, len(poly)):
        d.draw_line(poly[k - 1][0], poly[k - 1][1], poly[k][0], poly[k][1], pot[k])
    if d.n_outside:
        raise ValueError("%d boundary points fall outside the grid" % d.n_outside)
    d.fill()
    d.trim()
    return d.marks()


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
