"""Pure-Python restatement of spiffe's boundary file, boundary drawing and
metal marks (src/geometry.c:133-316, src/grid_drawing.c).

TEST INFRASTRUCTURE ONLY: the checker of the library's geometry entry points
(spf_geometry_parse / spf_geometry_draw); only tests/ load it.  Loops are
written out as the reference writes them, so keep grids small (a few hundred
points a side).  The field-flag bits are added by the C restatement
(oracle.derive_flags).
"""
import math
import re

DBL_MAX = 1.7976931348623157e308
FL_IS_SURFACE, FL_IS_INSIDE, FL_IS_VACUUM = 1, 2, 16
PI = 3.141592653589793238462643


def parse_points(text, zr_factor=1.0):
    """The &po/&point namelists of a boundary file, src/geometry.c:133-209.
    Returns a list of dicts with the fields of the reference's POINTSP."""
    text = re.sub(r"!.*", "", text).lower()
    # a group runs from "&name" to "&end" or to a bare "&"
    groups = re.findall(r"&(\w+)(.*?)&(?:end\b)?", text, flags=re.S)
    pts = []
    aovrb = 1.0                      # not reset between points
    for name, body in groups:
        if name not in ("po", "point"):
            raise ValueError("unrecognized namelist in boundary defintion")
        vals = dict((k, v) for k, v in re.findall(r"(\w+)\s*=\s*([^,\s]+)", body))
        nt = int(vals.get("nt", 1))
        x = float(vals["x"]) if "x" in vals else DBL_MAX
        y = float(vals["y"]) if "y" in vals else DBL_MAX
        theta = float(vals["theta"]) if "theta" in vals else DBL_MAX
        x0, y0 = float(vals.get("x0", 0)), float(vals.get("y0", 0))
        r, a, b = float(vals.get("r", 0)), float(vals.get("a", 0)), float(vals.get("b", 0))
        if "aovrb" in vals:
            aovrb = float(vals["aovrb"])
        pot = float(vals.get("potential", 0))
        mat, ramp, chg = int(vals.get("material_id", 0)), int(vals.get("ramp_potential", 0)), \
            int(vals.get("change_direction", 0))
        if theta != DBL_MAX:
            theta *= PI / 180.0
        if not pts and nt != 1:
            raise ValueError("first point in geometry definition must have nt=1")
        p = dict(nt=nt, x=x * zr_factor, y=y * zr_factor, x0=x0 * zr_factor, y0=y0 * zr_factor, r=r * zr_factor)
        a *= zr_factor
        b *= zr_factor
        if x != DBL_MAX:
            x *= zr_factor
        if y != DBL_MAX:
            y *= zr_factor
        if nt == 2:
            if theta == DBL_MAX and (x == DBL_MAX or y == DBL_MAX):
                raise ValueError("If NT=2, must give theta or (x, y)")
            if theta != DBL_MAX and (x != DBL_MAX or y != DBL_MAX):
                raise ValueError("If NT=2, must give theta or (x, y), not both")
            if r == 0:
                if a == 0 and b == 0:
                    raise ValueError("Can't have NT=2, R=0, A=0, and B=0")
                if a == 0 or b == 0:
                    if aovrb <= 0:
                        raise ValueError("Can't have NT=2, R=0, AOVRB<=0 and A=0 or B=0")
                    if a == 0:
                        a = b * aovrb
                    else:
                        b = a / aovrb
                if theta == DBL_MAX:
                    theta = math.atan2(y / b, x / a)
            elif theta == DBL_MAX:
                theta = math.atan2(y, x)
        if chg:
            theta = 2 * PI + theta
        p.update(a=a, b=b, potential=pot, material_id=mat, ramp_potential=ramp, theta=theta)
        pts.append(p)
    if not pts:
        raise ValueError("too few boundary points in geometry file")
    return pts


def _c_int(v):
    """C's (int) of a double: truncation toward zero."""
    return int(v) if math.isfinite(v) else -(1 << 31)


class Drawing:
    """The drawing of src/grid_drawing.c on an (nz+1) x (nr+1) array (points
    with i < nz, j < nr are drawn)."""

    def __init__(self, nz, nr, zmin, zmax, rmax):
        self.nx, self.ny = nz, nr
        self.xmin, self.xmax, self.ymin, self.ymax = zmin, zmax, 0.0, rmax
        self.dx = (zmax - zmin) / (nz - 1)
        self.dy = (rmax - 0.0) / (nr - 1)
        self.L = [[0] * (nr + 1) for _ in range(nz + 1)]
        self.P = [[0.0] * (nr + 1) for _ in range(nz + 1)]
        self.M = [[0] * (nr + 1) for _ in range(nz + 1)]
        self.last = None
        self.n_outside = 0
        self.vacuum = None

    def add_point(self, x, y, pot, mat):   # src/grid_drawing.c:806-845
        if self.last is not None and self.last == (x, y):
            return
        i = _c_int((x - self.xmin) / self.dx + 0.5)
        j = _c_int((y - self.ymin) / self.dy + 0.5)
        if 0 <= i < self.nx and 0 <= j < self.ny:
            self.L[i][j] = FL_IS_SURFACE
            self.P[i][j] = pot
            if self.M[i][j] == 0:
                self.M[i][j] = mat
        else:
            self.n_outside += 1
        self.last = (x, y)

    def draw_line(self, xi, yi, xf, yf, pot, mat, start_pot):   # src/grid_drawing.c:215-295
        dx, dy = self.dx, self.dy
        if abs(xi - xf) >= dx:
            slope = (yf - yi) / (xf - xi)
            if abs(yf - yi) < dy / 1e6:
                slope = 0.0
            intercept = yi - slope * xi
            step = dx
            if abs(slope * step) > abs(dy):
                step = dy * .5 / abs(slope)
            pslope = (pot - start_pot) / (xf - xi)
            step = 0.5 * step
            x = xi
            while True:
                self.add_point(x, slope * x + intercept, start_pot + pslope * (x - xi), mat)
                if xi <= xf:
                    x += step
                    if not x <= xf:
                        break
                else:
                    x -= step
                    if not x >= xf:
                        break
            self.add_point(xf, yf, pot, mat)
        elif abs(yi - yf) >= dy:
            slope = (xf - xi) / (yf - yi)
            if abs(xf - xi) < dx / 1e6:
                slope = 0.0
            intercept = xi - slope * yi
            pslope = (pot - start_pot) / (yf - yi)
            step = dy * 0.5
            y = yi
            while True:
                self.add_point(slope * y + intercept, y, start_pot + pslope * (y - yi), mat)
                if yi <= yf:
                    y += step
                    if not y <= yf:
                        break
                else:
                    y -= step
                    if not y >= yf:
                        break
            self.add_point(xf, yf, pot, mat)
        else:
            self.add_point(xf, yf, pot, mat)

    def draw_arc(self, xi, yi, xc, yc, a, b, circle, thetaf, pot, mat):   # src/grid_drawing.c:297-402
        dx, dy = self.dx, self.dy
        if circle:
            theta = math.atan2(yi - yc, xi - xc)
            dtheta = min(dx, dy) / a / 2
        else:
            theta = math.atan2((yi - yc) / b, (xi - xc) / a)
            dtheta = min(dx / a, dy / b) / 2
        sense = theta < thetaf
        if sense:
            step = dtheta
            while theta > thetaf:
                theta -= 2 * PI
        else:
            while theta < thetaf:
                theta += 2 * PI
            step = -dtheta
        n = _c_int(abs((thetaf - theta) / step))
        if n < 15:
            sgn = 1 if step > 0 else (-1 if step < 0 else 0)
            n = 15
            step = sgn * abs(thetaf - theta) / n
        step *= 0.5
        n *= 2
        theta0 = theta
        for i in range(1, n):
            self.add_point(xc + a * math.cos(theta), yc + b * math.sin(theta), pot, mat)
            theta = theta0 + i * step
        self.add_point(xc + a * math.cos((thetaf + theta) / 2), yc + b * math.sin((thetaf + theta) / 2), pot, mat)
        x, y = xc + a * math.cos(thetaf), yc + b * math.sin(thetaf)
        self.add_point(x, y, pot, mat)
        return x, y

    def fill(self):   # src/grid_drawing.c:889-1084, the sweeps as written
        L, nx, ny = self.L, self.nx, self.ny
        if self.vacuum is None:
            for i in range(nx):
                if L[i][0] and not L[i][1]:
                    break
            else:
                raise ValueError("can't determine what points are in vacuum")
            if i > 0:
                L[i][1] = FL_IS_VACUUM
        else:
            L[self.vacuum[0]][self.vacuum[1]] = FL_IS_VACUUM

        def visit(i, j):
            n = 0
            if L[i][j] == FL_IS_VACUUM:
                for a, b in ((i, j + 1), (i, j - 1), (i + 1, j), (i - 1, j)):
                    if L[a][b] == 0:
                        L[a][b] = FL_IS_VACUUM
                        n += 1
            return n

        changed = 1
        while changed:
            changed = 0
            for i in range(1, nx - 1):
                for j in range(1, ny - 1):
                    changed += visit(i, j)
                for j in range(ny - 2, 0, -1):
                    changed += visit(i, j)
            for i in range(nx - 2, 0, -1):
                for j in range(1, ny - 1):
                    changed += visit(i, j)
                for j in range(ny - 2, 0, -1):
                    changed += visit(i, j)
        for i in range(nx):
            for j in range(ny):
                if L[i][j] == 0:
                    L[i][j] = FL_IS_INSIDE
                elif L[i][j] == FL_IS_VACUUM:
                    L[i][j] = 0
        for i in range(1, nx - 1):
            for j in range(1, ny - 1):
                if L[i][j] != FL_IS_INSIDE:
                    continue
                change, psum, nsum = 0, 0.0, 0
                for i1 in (-1, 0, 1):
                    for j1 in (-1, 0, 1):
                        if i1 == 0 and j1 == 0:
                            continue
                        v = L[i + i1][j + j1]
                        if not v:
                            change += 1
                        elif v == FL_IS_SURFACE:
                            psum += self.P[i + i1][j + j1]
                            nsum += 1
                if change == 8:
                    L[i][j] = 0
                elif change:
                    L[i][j] = FL_IS_SURFACE
                    self.P[i][j] = psum / nsum if nsum else 0.0

    def trim(self):   # src/grid_drawing.c:1165-1192
        L, nx, ny = self.L, self.nx, self.ny
        for j in range(ny):
            if L[0][j] and not L[1][j]:
                L[0][j] = 0
        for j in range(ny):
            if L[nx - 1][j] and not L[nx - 2][j]:
                L[nx - 1][j] = 0
        for i in range(nx):
            if L[i][0] and not L[i][1]:
                L[i][0] = 0


def draw(nz, nr, zmin, zmax, rmax, pts):
    """src/geometry.c:211-316: marks (surface/inside), potentials and material ids
    as nested lists [iz][ir] of size (nz+1) x (nr+1)."""
    D = Drawing(nz, nr, zmin, zmax, rmax)
    xi, yi = pts[0]["x"] + pts[0]["x0"], pts[0]["y"] + pts[0]["y0"]
    D.add_point(xi, yi, pts[0]["potential"], pts[0]["material_id"])
    for k in range(1, len(pts)):
        p = pts[k]
        if p["nt"] == 1:
            xf, yf = p["x"] + p["x0"], p["y"] + p["y0"]
            start = pts[k - 1]["potential"] if p["ramp_potential"] else p["potential"]
            D.draw_line(xi, yi, xf, yf, p["potential"], p["material_id"], start)
            if D.n_outside:
                raise ValueError("boundary points outside the simulation region")
            xi, yi = xf, yf
        elif p["nt"] == 2:
            if p["r"] != 0:
                xi, yi = D.draw_arc(xi, yi, p["x0"], p["y0"], p["r"], p["r"], True, p["theta"], p["potential"],
                                    p["material_id"])
            else:
                xi, yi = D.draw_arc(xi, yi, p["x0"], p["y0"], p["a"], p["b"], False, p["theta"], p["potential"],
                                    p["material_id"])
            if D.n_outside:
                raise ValueError("boundary points outside the simulation region")
        elif p["nt"] == 3:
            xi, yi = p["x"] + p["x0"], p["y"] + p["y0"]
            D.add_point(xi, yi, p["potential"], p["material_id"])
        elif p["nt"] == 4:
            x, y = p["x"] + p["x0"], p["y"] + p["y0"]
            i, j = _c_int((x - D.xmin) / D.dx + 0.5), _c_int((y - D.ymin) / D.dy + 0.5)
            if not (0 <= i < D.nx and 0 <= j < D.ny):
                raise ValueError("vacuum point out of bounds")
            D.vacuum = (i, j)
    D.fill()
    D.trim()
    return D.L, D.P, D.M
