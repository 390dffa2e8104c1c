"""Boundary files, drawing and metal flags: the library's C++ restatement of
src/geometry.c and src/grid_drawing.c against the pure-Python oracle (bit for
bit), and the oracle against closed forms.  Host code only: no GPU needed.

tests/golden/*.geo are the reference's example boundary files
(examples/dcGun/dc.geo, examples/apsTCRFGun/{e60,excite}/mg6mm-*.geo), kept
as input fixtures."""
import math
import os

import numpy as np
import pytest

from oracle import geometry_ref as G
from oracle import oracle as O
from spiffe_amd import abi
from spiffe_amd import geometry as PG

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
DECKS = [("dc.geo", 165, 165, 0.0, 0.072, 0.042),                       # examples/dcGun/dc.spiffe
         ("mg6mm-3.geo", 165, 165, 0.0, 0.0959148653920, 0.04289777777),  # excite/mg6mm-3.spiffe
         ("mg6mm-e60-j0.geo", 165, 165, 0.0, 0.0959148653920, 0.04289777777)]


def read(name):
    with open(os.path.join(GOLD, name)) as fp:
        return fp.read()


def oracle_draw(nz, nr, zmin, zmax, rmax, text, zr=1.0):
    L, P, M = G.draw(nz, nr, zmin, zmax, rmax, G.parse_points(text, zr))
    return O.derive_flags(np.array(L, dtype=np.int16)), np.array(P), np.array(M, dtype=np.int16)


def assert_same(a, b):
    fo, po, mo = a
    fd, pd, md = b
    assert np.array_equal(fo, fd), "flags differ at %d points" % np.count_nonzero(fo != fd)
    assert np.array_equal(po.view(np.int64), pd.view(np.int64)), "potentials differ"
    assert np.array_equal(mo, md), "material ids differ"


@pytest.mark.parametrize("name", [d[0] for d in DECKS])
def test_boundary_files_parse_alike(name):
    text = read(name)
    po, pd = G.parse_points(text), PG.parse_boundary(text)
    assert len(po) == len(pd)
    for a, b in zip(po, pd):
        for k in ["nt", "x", "y", "x0", "y0", "r", "theta", "a", "b", "potential", "material_id", "ramp_potential"]:
            assert a[k] == getattr(b, k), (k, a[k], getattr(b, k))


def test_dc_gun_boundary_values():
    pts = PG.parse_boundary(read("dc.geo"))
    assert len(pts) == 14
    arc = pts[4]
    assert (arc.nt, arc.x0, arc.y0, arc.r) == (2, 0.009, 0.028, 0.008)
    assert arc.theta == 90 * (math.pi / 180.0)
    assert pts[8].potential == 1e6 and pts[0].potential == 0
    assert pts[1].nt == 1 and pts[1].y == 0.01
    scaled = PG.parse_boundary(read("dc.geo"), zr_factor=2.0)
    assert scaled[4].x0 == 0.018 and scaled[4].r == 0.016 and scaled[4].theta == arc.theta


@pytest.mark.parametrize("deck", DECKS, ids=[d[0] for d in DECKS])
def test_example_geometries_match_the_oracle(deck):
    name, nz, nr, zmin, zmax, rmax = deck
    text = read(name)
    assert_same(oracle_draw(nz, nr, zmin, zmax, rmax, text), PG.draw(nz, nr, zmin, zmax, rmax, PG.parse_boundary(text)))


@pytest.mark.parametrize("nz,nr", [(61, 41), (97, 53), (210, 95)])
def test_example_geometries_on_other_meshes(nz, nr):
    text = read("mg6mm-e60-j0.geo")
    assert_same(oracle_draw(nz, nr, 0.0, 0.0959148653920, 0.04289777777, text),
                PG.draw(nz, nr, 0.0, 0.0959148653920, 0.04289777777, PG.parse_boundary(text)))


SHAPES = {
    # two shapes (nt=3), an ellipse arc, a ramped potential and material ids, an explicit vacuum point
    "multi": """
 &po x=0.0, y=0.0 &end
 &po x=0.0, y=0.02, potential=5, material_id=2 &end
 &po x=0.03, y=0.02, potential=5 &end
 &po nt=2, x0=0.03, y0=0.012, a=0.006, b=0.008, theta=-90, potential=5 &end
 &po x=0.05, y=0.004, potential=-3, ramp_potential=1, material_id=3 &end
 &po x=0.05, y=0.0 &end
 &po x=0.0, y=0.0 &end
 &po nt=3, x=0.012, y=0.006, potential=9, material_id=4 &end
 &po x=0.016, y=0.006, potential=9 &end
 &po x=0.016, y=0.009, potential=9 &end
 &po x=0.012, y=0.009, potential=9 &end
 &po x=0.012, y=0.006, potential=9 &end
 &po nt=4, x=0.02, y=0.002 &end
""",
    # an ellipse given by aovrb and an end point, a circle drawn the other way (change_direction)
    "arcs": """
 &po x=0.0, y=0.0 &
 &po x=0.0, y=0.01 &
 &po nt=2, x0=0.015, y0=0.01, b=0.01, aovrb=1.5, x=0.0, y=0.01 &
 &po x=0.035, y=0.02 &
 &po nt=2, x0=0.045, y0=0.02, r=0.01, x=0.0, y=-0.01, change_direction=1 &
 &po x=0.055, y=0.0 &
 &po x=0.0, y=0.0 &
""",
}


@pytest.mark.parametrize("shape", sorted(SHAPES))
def test_drawn_shapes_match_the_oracle(shape):
    text = SHAPES[shape]
    for nz, nr in [(71, 31), (150, 64)]:
        assert_same(oracle_draw(nz, nr, 0.0, 0.06, 0.03, text),
                    PG.draw(nz, nr, 0.0, 0.06, 0.03, PG.parse_boundary(text)))


# ---- the oracle against closed forms ----

def test_oracle_rectangle_on_grid_points():
    # a box with vertices on grid points: its edges are exactly the drawn
    # surface, the points strictly inside are vacuum, the rest is metal; the
    # box's axis edge, with vacuum above it, is trimmed (src/grid_drawing.c:1187-1189)
    nz, nr, L, R = 41, 21, 0.04, 0.02
    dz, dr = L / (nz - 1), R / (nr - 1)
    z1, z2, r2 = 5 * dz, 30 * dz, 12 * dr
    text = "&po x=%r, y=0 &end &po x=%r, y=%r &end &po x=%r, y=%r &end &po x=%r, y=0 &end &po x=%r, y=0 &end" % (
        z1, z1, r2, z2, r2, z2, z1)
    f, _, _ = oracle_draw(nz, nr, 0.0, L, R, text)
    for i in range(nz):
        for j in range(nr):
            m = f[i, j] & 3
            on_edge = (i in (5, 30) and 0 <= j <= 12) or (j == 12 and 5 <= i <= 30)
            if 5 < i < 30 and j < 12:
                assert m == 0, (i, j)
            elif on_edge:
                assert m == abi.FL_IS_SURFACE, (i, j)
            else:
                assert m == abi.FL_IS_INSIDE, (i, j)


def point_in_polygon(x, y, poly):
    inside = False
    for (x1, y1), (x2, y2) in zip(poly, poly[1:] + poly[:1]):
        if (y1 > y) != (y2 > y) and x < x1 + (y - y1) * (x2 - x1) / (y2 - y1):
            inside = not inside
    return inside


def dist_to_polygon(x, y, poly):
    d = 1e300
    for (x1, y1), (x2, y2) in zip(poly, poly[1:] + poly[:1]):
        vx, vy = x2 - x1, y2 - y1
        t = max(0.0, min(1.0, ((x - x1) * vx + (y - y1) * vy) / (vx * vx + vy * vy or 1e-300)))
        d = min(d, math.hypot(x - x1 - t * vx, y - y1 - t * vy))
    return d


def test_oracle_polygon_agrees_with_point_in_polygon_away_from_the_boundary():
    poly = [(0.0, 0.0), (0.0, 0.042), (0.030, 0.042), (0.030, 0.0125), (0.035, 0.0125), (0.035, 0.042),
            (0.105, 0.042), (0.105, 0.010), (0.1395, 0.010), (0.1395, 0.0)]
    nz, nr, L, R = 141, 61, 0.14, 0.06
    dz, dr = L / (nz - 1), R / (nr - 1)
    text = "".join("&po x=%r, y=%r &end\n" % p for p in poly + [poly[0]])
    f, _, _ = oracle_draw(nz, nr, 0.0, L, R, text)
    checked = 0
    for i in range(nz):
        for j in range(1, nr):
            z, r = i * dz, j * dr
            if dist_to_polygon(z, r, poly) < 1.5 * max(dz, dr):
                continue
            metal = (f[i, j] & 3) != 0
            if i == nz - 1 or j == nr - 1 or i == 0:
                continue
            assert metal == (not point_in_polygon(z, r, poly)), (i, j)
            checked += 1
    assert checked > 3000


def test_oracle_circle_cavity():
    # a circular cavity of radius R0 centred on the axis, closed by the axis
    nz, nr, L, R = 121, 61, 0.06, 0.03
    R0, zc = 0.025, 0.03
    text = ("&po x=%r, y=0 &end\n&po nt=2, x0=%r, y0=0, r=%r, theta=0 &end\n&po x=%r, y=0 &end"
            % (zc - R0, zc, R0, zc - R0))
    f, _, _ = oracle_draw(nz, nr, 0.0, L, R, text)
    dz, dr = L / (nz - 1), R / (nr - 1)
    for i in range(1, nz - 1):
        for j in range(1, nr - 1):
            d = math.hypot(i * dz - zc, j * dr)
            if d < R0 - 1.5 * dz:
                assert f[i, j] & 3 == 0, (i, j)
            elif d > R0 + 1.5 * dz:
                assert f[i, j] & 3 != 0, (i, j)
            else:
                continue
    # the drawn surface points lie within half a cell of the circle
    surf = np.argwhere(f & abi.FL_IS_SURFACE)
    for i, j in surf:
        d = math.hypot(i * dz - zc, j * dr)
        if j > 0:
            assert abs(d - R0) < 1.5 * dz


def test_oracle_ramped_potential_is_linear_along_the_line():
    nz, nr, L, R = 81, 41, 0.08, 0.04
    text = ("&po x=0.0, y=0.0 &end &po x=0.0, y=0.03, potential=0 &end "
            "&po x=0.07, y=0.03, potential=7000, ramp_potential=1 &end &po x=0.07, y=0.0, potential=7000 &end "
            "&po x=0.0, y=0.0 &end")
    _, P, _ = oracle_draw(nz, nr, 0.0, L, R, text)
    dz = L / (nz - 1)
    j = int(0.03 / (R / (nr - 1)) + 0.5)
    for i in range(1, 70):
        z = i * dz
        if 0.001 < z < 0.069:
            assert P[i, j] == pytest.approx(7000 * z / 0.07, abs=7000 * dz / 0.07)


@pytest.mark.parametrize("text,msg", [
    ("&po nt=2, x0=0, y0=0, r=0.01, theta=0 &end", "first point"),
    ("&po x=0, y=0 &end &po nt=2, x0=0.01, y0=0, r=0.01 &end", "must give theta"),
    ("&po x=0, y=0 &end &po nt=2, x0=0.01, y0=0, theta=10 &end", "R=0, A=0, and B=0"),
    ("&po x=0, y=0 &end &define_cathode z_position=0 &end", "unrecognized namelist"),
    ("&po x=0, y=0, colour=3 &end", "unknown point parameter"),
    ("&po x=0, y=0", "has no &end"),
])
def test_boundary_file_errors(text, msg):
    with pytest.raises(ValueError, match=msg):
        PG.parse_boundary(text)


def test_boundary_outside_the_grid_is_an_error():
    with pytest.raises(ValueError, match="outside the simulation region"):
        PG.draw(41, 21, 0.0, 0.04, 0.02, PG.parse_boundary("&po x=0, y=0 &end &po x=0.05, y=0.01 &end"))
