// spf_geometry.cpp -- the boundary file, the boundary drawing and the metal
// flags: what process_geometry_definition (src/geometry.c:21-369) and the
// drawing routines of src/grid_drawing.c leave in FIELDS.metal_flags,
// PsiImposed and materialID.  Host code, no GPU: the flags go to the device
// through spf_run_create, the potentials through spf_field_upload.
// Build with -ffp-contract=off and without sin/cos builtins (the arcs are
// drawn with libm's sin, cos and atan2 exactly as written).
#include <float.h>
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <deque>
#include <string>
#include <vector>

#include "../../include/spiffe_amd.h"
#include "spf_namelist.h"

namespace {

const double kPi = 3.141592653589793238462643;

int fail(char *err, int64_t len, const std::string &msg) {
  if (err && len > 0) snprintf(err, (size_t)len, "%s", msg.c_str());
  return SPF_EINVAL;
}

// C conversion of a double to int: truncation toward zero.  `ok` is false
// when the value is out of [lo, hi) after truncation or not a number.
int trunc_index(double v, int hi, bool *ok) {
  *ok = v > -1.0 && v < (double)hi;
  return *ok ? (int)v : -1;
}

// The drawing state of src/grid_drawing.c:30-48 for one call.
struct Drawing {
  int nx, ny;
  double xmin, xmax, ymin, ymax, dx, dy;
  int ldr;                      // row length of the (nz+1) x (nr+1) arrays
  int16_t *lat;
  double *pot;
  int16_t *mat;
  bool have_last = false;
  double last_x = 0, last_y = 0;
  int64_t n_outside = 0;
  bool vacuum_defined = false;
  int ixv = 0, iyv = 0;

  int16_t &L(int i, int j) { return lat[(int64_t)i * ldr + j]; }
  double &P(int i, int j) { return pot[(int64_t)i * ldr + j]; }
  int16_t &M(int i, int j) { return mat[(int64_t)i * ldr + j]; }

  // add_point, src/grid_drawing.c:806-845
  void add_point(double x, double y, double potential0, int16_t material0) {
    if (have_last && last_x == x && last_y == y) return;
    bool oki, okj;
    int i = trunc_index((x - xmin) / dx + 0.5, nx, &oki);
    int j = trunc_index((y - ymin) / dy + 0.5, ny, &okj);
    if (oki && okj) {
      L(i, j) = SPF_FL_IS_SURFACE;
      P(i, j) = potential0;
      if (M(i, j) == 0) M(i, j) = material0;
    } else {
      n_outside++;
    }
    have_last = true;
    last_x = x;
    last_y = y;
  }

  // draw_line, src/grid_drawing.c:215-295
  void draw_line(double xi, double yi, double xf, double yf, double potential, int16_t material,
                 double startPotential) {
    double slope, intercept, x, y, _dx, _dy, potSlope;
    if (fabs(xi - xf) >= dx) {
      slope = (yf - yi) / (xf - xi);
      if (fabs(yf - yi) < dy / 1e6) slope = 0;
      intercept = yi - slope * xi;
      _dx = dx;
      if (fabs(slope * _dx) > fabs(dy)) _dx = dy * .5 / fabs(slope);
      potSlope = (potential - startPotential) / (xf - xi);
      _dx = 0.5 * _dx;
      x = xi;
      if (xi <= xf) {
        do {
          y = slope * x + intercept;
          add_point(x, y, startPotential + potSlope * (x - xi), material);
        } while ((x += _dx) <= xf);
      } else {
        do {
          y = slope * x + intercept;
          add_point(x, y, startPotential + potSlope * (x - xi), material);
        } while ((x -= _dx) >= xf);
      }
      add_point(xf, yf, potential, material);
    } else if (fabs(yi - yf) >= dy) {
      slope = (xf - xi) / (yf - yi);
      if (fabs(xf - xi) < dx / 1e6) slope = 0;
      intercept = xi - slope * yi;
      potSlope = (potential - startPotential) / (yf - yi);
      _dy = dy * 0.5;
      y = yi;
      if (yi <= yf) {
        do {
          x = slope * y + intercept;
          add_point(x, y, startPotential + potSlope * (y - yi), material);
        } while ((y += _dy) <= yf);
      } else {
        do {
          x = slope * y + intercept;
          add_point(x, y, startPotential + potSlope * (y - yi), material);
        } while ((y -= _dy) >= yf);
      }
      add_point(xf, yf, potential, material);
    } else {
      add_point(xf, yf, potential, material);
    }
  }

  // draw_circle and draw_ellipse, src/grid_drawing.c:297-402: an arc from the
  // current point to angle thetaf, in steps of at most half a cell (and at
  // least 30 points); a == b == r for a circle
  void draw_arc(double xi, double yi, double xc, double yc, double a, double b, bool circle, double thetaf,
                double *xf, double *yf, double potential, int16_t material) {
    double theta, x, y, _dtheta, theta0, dtheta;
    int i, n, sense;
    if (circle)
      theta = atan2(yi - yc, xi - xc);
    else
      theta = atan2((yi - yc) / b, (xi - xc) / a);
    sense = theta < thetaf ? 1 : 0;
    if (circle) {
      if (a != 0)
        dtheta = (dx < dy ? dx : dy) / a / 2;
      else if (dx < dy)
        dtheta = dx / (xmax - xmin) / 2;
      else
        dtheta = dy / (ymax - ymin) / 2;
    } else {
      double u = dx / a, v = dy / b;
      dtheta = (u < v ? u : v) / 2;
    }
    if (sense) {
      _dtheta = dtheta;
      while (theta > thetaf) theta -= 2 * kPi;
    } else {
      while (theta < thetaf) theta += 2 * kPi;
      _dtheta = -dtheta;
    }
    if ((n = (int)fabs((thetaf - theta) / _dtheta)) < 15) {
      int sgn = _dtheta > 0 ? 1 : (_dtheta < 0 ? -1 : 0);
      _dtheta = sgn * fabs(thetaf - theta) / (n = 15);
    }
    _dtheta *= 0.5;
    n *= 2;
    theta0 = theta;
    for (i = 1; i < n; i++) {
      x = xc + a * cos(theta);
      y = yc + b * sin(theta);
      add_point(x, y, potential, material);
      theta = theta0 + i * _dtheta;
    }
    x = xc + a * cos((thetaf + theta) / 2);
    y = yc + b * sin((thetaf + theta) / 2);
    add_point(x, y, potential, material);
    x = xc + a * cos(thetaf);
    y = yc + b * sin(thetaf);
    add_point(x, y, potential, material);
    *xf = x;
    *yf = y;
  }

  // set_vacuum_point, src/grid_drawing.c:790-800
  void set_vacuum_point(double x, double y) {
    bool oki, okj;
    double vi = (x - xmin) / dx + 0.5, vj = (y - ymin) / dy + 0.5;
    ixv = trunc_index(vi, nx, &oki);
    iyv = trunc_index(vj, ny, &okj);
    vacuum_defined = true;
    if (!oki || !okj) n_outside++;
  }

  // fill_in_drawing, src/grid_drawing.c:889-1084.  The reference's repeated
  // sweeps reach the closure of the seed: every zero point joined to it by
  // 4-neighbour steps taken from points off the lattice edges.  A queue
  // reaches the same set.
  int fill(char *err, int64_t errlen) {
    const int16_t VAC = SPF_FL_IS_VACUUM;
    std::deque<std::pair<int, int>> q;
    if (!vacuum_defined) {
      int i;
      for (i = 0; i < nx; i++)
        if (L(i, 0) && !L(i, 1)) break;
      if (i == nx)
        return fail(err, errlen,
                    "can't determine what points are in vacuum--use nt=4 point to mark a point that is in vacuum");
      if (i > 0) {
        L(i, 1) = VAC;
        q.push_back({i, 1});
      }
    } else {
      if (ixv < 0 || iyv < 0) return fail(err, errlen, "vacuum point out of bounds");
      L(ixv, iyv) = VAC;
      q.push_back({ixv, iyv});
    }
    while (!q.empty()) {
      auto [i, j] = q.front();
      q.pop_front();
      if (i < 1 || i > nx - 2 || j < 1 || j > ny - 2) continue;   // edge points do not spread
      const int di[4] = {0, 0, 1, -1}, dj[4] = {1, -1, 0, 0};
      for (int k = 0; k < 4; k++) {
        int a = i + di[k], b = j + dj[k];
        if (L(a, b) == 0) {
          L(a, b) = VAC;
          q.push_back({a, b});
        }
      }
    }
    for (int i = 0; i < nx; i++)
      for (int j = 0; j < ny; j++) {
        if (L(i, j) == 0)
          L(i, j) = SPF_FL_IS_INSIDE;
        else if (L(i, j) == VAC)
          L(i, j) = 0;
      }
    // inside points next to vacuum become surface, one pass in index order
    for (int i = 1; i < nx - 1; i++)
      for (int j = 1; j < ny - 1; j++) {
        if (L(i, j) != SPF_FL_IS_INSIDE) continue;
        int change = 0;
        double potentialSum = 0;
        long nSummed = 0;
        for (int i1 = -1; i1 <= 1; i1++)
          for (int j1 = -1; j1 <= 1; j1++) {
            if (i1 == 0 && j1 == 0) continue;
            int16_t v = L(i + i1, j + j1);
            if (!v) {
              change++;
            } else if (v == SPF_FL_IS_SURFACE) {
              potentialSum += P(i + i1, j + j1);
              nSummed++;
            }
          }
        if (change == 8) {
          L(i, j) = 0;
        } else if (change) {
          L(i, j) = SPF_FL_IS_SURFACE;
          P(i, j) = nSummed ? potentialSum / nSummed : 0;
        }
      }
    return 0;
  }

  // trim_drawing, src/grid_drawing.c:1165-1192
  void trim() {
    for (int j = 0; j < ny; j++)
      if (L(0, j) && !L(1, j)) L(0, j) = 0;
    int i = nx - 1;
    for (int j = 0; j < ny; j++)
      if (L(i, j) && !L(i - 1, j)) L(i, j) = 0;
    for (i = 0; i < nx; i++)
      if (L(i, 0) && !L(i, 1)) L(i, 0) = 0;
  }
};

// The field flags of src/geometry.c:317-362 from the surface/inside marks.
void derive_flags(int16_t *f, int nz, int nr) {
  const int ld = nr + 1;
  auto F = [&](int iz, int ir) -> int16_t & { return f[(int64_t)iz * ld + ir]; };
  for (int iz = 0; iz < nz + 1; iz++)
    for (int ir = 0; ir < nr + 1; ir++) {
      if (!(F(iz, ir) & (SPF_FL_IS_SURFACE | SPF_FL_IS_INSIDE))) continue;
      const int M = SPF_FL_IS_SURFACE | SPF_FL_IS_INSIDE;
      F(iz, ir) |= SPF_FL_EPHI_ZERO;
      bool up = iz && (F(iz - 1, ir) & M);
      if (up) F(iz, ir) |= SPF_FL_EZ_ZERO | SPF_FL_BR_ZERO;
      if (ir == 0 || (F(iz, ir - 1) & M)) F(iz, ir) |= SPF_FL_ER_ZERO | SPF_FL_BZ_ZERO;
      if (up && (ir == 0 || ((F(iz, ir - 1) & M) && (F(iz - 1, ir - 1) & M)))) F(iz, ir) |= SPF_FL_BPHI_ZERO;
    }
  for (int ir = 0; ir < nr; ir++) {
    int iz = nz - 1;
    while (iz >= 0 && (F(iz, ir) & SPF_FL_IS_INSIDE)) iz--;
    if (iz != 0 && (iz += 1) < nz - 1) F(iz, ir) |= SPF_FL_LAST_ACTIVE_PT;
  }
}

}  // namespace

extern "C" {

int spf_geometry_parse(const char *text, double zr_factor, spf_boundary_point *out, int64_t max_points,
                       int64_t *n_points, char *err, int64_t err_len) {
  if (!text || !n_points) return fail(err, err_len, "null argument");
  std::vector<spf::Namelist> groups;
  std::string e;
  if (spf::parse_namelists(text, groups, e)) return fail(err, err_len, e);
  static const char *const known[] = {"nt", "x", "y", "x0", "y0", "r", "theta", "a", "b", "aovrb", "potential",
                                      "material_id", "ramp_potential", "change_direction", nullptr};
  double aovrb = 1;   // keeps its value from one point to the next, as in the reference
  int64_t n = 0;
  for (const auto &g : groups) {
    if (g.group != "point" && g.group != "po") return fail(err, err_len, "unrecognized namelist in boundary defintion");
    std::string bad = g.unknown_keys(known);
    if (!bad.empty()) return fail(err, err_len, "unknown point parameter(s): " + bad);
    // per-point defaults, src/geometry.c:142-145
    int32_t nt = 1;
    int16_t change_direction = 0, material_id = 0, ramp_potential = 0;
    double x0 = 0, y0 = 0, r = 0, potential = 0, a = 0, b = 0;
    double x = DBL_MAX, y = DBL_MAX, theta = DBL_MAX;
    g.get("nt", &nt, e);
    g.get("x", &x, e);
    g.get("y", &y, e);
    g.get("x0", &x0, e);
    g.get("y0", &y0, e);
    g.get("r", &r, e);
    g.get("theta", &theta, e);
    g.get("a", &a, e);
    g.get("b", &b, e);
    g.get("aovrb", &aovrb, e);
    g.get("potential", &potential, e);
    g.get("material_id", &material_id, e);
    g.get("ramp_potential", &ramp_potential, e);
    g.get("change_direction", &change_direction, e);
    if (!e.empty()) return fail(err, err_len, e);
    if (theta != DBL_MAX) theta *= kPi / 180.0;
    if (n == 0 && nt != 1) return fail(err, err_len, "first point in geometry definition must have nt=1");
    spf_boundary_point p;
    memset(&p, 0, sizeof(p));
    p.nt = nt;
    p.x = x * zr_factor;
    p.y = y * zr_factor;
    p.x0 = x0 * zr_factor;
    p.y0 = y0 * zr_factor;
    p.r = r * zr_factor;
    a *= zr_factor;
    b *= zr_factor;
    if (x != DBL_MAX) x *= zr_factor;
    if (y != DBL_MAX) y *= zr_factor;
    if (nt == 2) {
      if (theta == DBL_MAX && (x == DBL_MAX || y == DBL_MAX)) return fail(err, err_len, "If NT=2, must give theta or (x, y)");
      if (theta != DBL_MAX && (x != DBL_MAX || y != DBL_MAX))
        return fail(err, err_len, "If NT=2, must give theta or (x, y), not both");
      if (r == 0) {
        if (a == 0 && b == 0) return fail(err, err_len, "Can't have NT=2, R=0, A=0, and B=0");
        if (a == 0 || b == 0) {
          if (aovrb <= 0) return fail(err, err_len, "Can't have NT=2, R=0, AOVRB<=0 and A=0 or B=0");
          if (a == 0)
            a = b * aovrb;
          else
            b = a / aovrb;
        }
        if (theta == DBL_MAX) theta = atan2(y / b, x / a);
      } else if (theta == DBL_MAX) {
        theta = atan2(y, x);
      }
    }
    if (change_direction) theta = 2 * kPi + theta;
    p.a = a;
    p.b = b;
    p.potential = potential;
    p.material_id = material_id;
    p.ramp_potential = ramp_potential;
    p.theta = theta;
    if (out && n < max_points) out[n] = p;
    n++;
  }
  if (n == 0) return fail(err, err_len, "too few boundary points in geometry file");
  *n_points = n;
  if (out && n > max_points) return fail(err, err_len, "more boundary points than max_points");
  return SPF_OK;
}

int spf_geometry_draw(int32_t nz, int32_t nr, double zmin, double zmax, double rmax, const spf_boundary_point *pnt,
                      int64_t n_points, int16_t *flags, double *psi_imposed, int16_t *material, char *err,
                      int64_t err_len) {
  if (nz < 2 || nr < 2) return fail(err, err_len, "nz and nr must be greater than 2");
  if (zmax <= zmin) return fail(err, err_len, "zmax must be greater than zmin");
  if (rmax <= 0) return fail(err, err_len, "rmax must be greater than 0");
  if (!pnt || n_points <= 0 || !flags) return fail(err, err_len, "too few boundary points in geometry file");
  if (pnt[0].nt != 1) return fail(err, err_len, "first point in geometry definition must have nt=1");
  const size_t cells = (size_t)(nz + 1) * (nr + 1);
  std::vector<double> pot_local;
  std::vector<int16_t> mat_local;
  if (!psi_imposed) {
    pot_local.assign(cells, 0.0);
    psi_imposed = pot_local.data();
  }
  if (!material) {
    mat_local.assign(cells, 0);
    material = mat_local.data();
  }
  memset(flags, 0, sizeof(int16_t) * cells);
  memset(psi_imposed, 0, sizeof(double) * cells);
  memset(material, 0, sizeof(int16_t) * cells);
  // set_up_drawing, src/grid_drawing.c:159-199
  Drawing D;
  D.nx = nz;
  D.ny = nr;
  D.xmin = zmin;
  D.xmax = zmax;
  D.ymin = 0.0;
  D.ymax = rmax;
  D.dx = (D.xmax - D.xmin) / (D.nx - 1);
  D.dy = (D.ymax - D.ymin) / (D.ny - 1);
  D.ldr = nr + 1;
  D.lat = flags;
  D.pot = psi_imposed;
  D.mat = material;
  // src/geometry.c:215-290
  double xi = pnt[0].x + pnt[0].x0, yi = pnt[0].y + pnt[0].y0, xf, yf;
  D.add_point(xi, yi, pnt[0].potential, pnt[0].material_id);
  char msg[256];
  for (int64_t k = 1; k < n_points; k++) {
    const spf_boundary_point &p = pnt[k];
    switch (p.nt) {
      case 1:
        xf = p.x + p.x0;
        yf = p.y + p.y0;
        D.draw_line(xi, yi, xf, yf, p.potential, p.material_id, p.ramp_potential ? pnt[k - 1].potential : p.potential);
        if (D.n_outside) {
          snprintf(msg, sizeof(msg),
                   "points on the boundary that were outside the simulation region for line: xi = %e, yi=%e, "
                   "xf=%e, yf=%e",
                   xi, yi, xf, yf);
          return fail(err, err_len, msg);
        }
        xi = xf;
        yi = yf;
        break;
      case 2:
        if (p.r != 0)
          D.draw_arc(xi, yi, p.x0, p.y0, p.r, p.r, true, p.theta, &xf, &yf, p.potential, p.material_id);
        else
          D.draw_arc(xi, yi, p.x0, p.y0, p.a, p.b, false, p.theta, &xf, &yf, p.potential, p.material_id);
        xi = xf;
        yi = yf;
        if (D.n_outside) {
          snprintf(msg, sizeof(msg),
                   "points on the boundary that were outside the simulation region for circle/ellipse: x0 = %e, "
                   "y0=%e, r=%e, a=%e, b=%e, theta=%e",
                   p.x0, p.y0, p.r, p.a, p.b, p.theta * 180 / kPi);
          return fail(err, err_len, msg);
        }
        break;
      case 3:
        xi = p.x + p.x0;
        yi = p.y + p.y0;
        D.add_point(xi, yi, p.potential, p.material_id);   // add_breakpoint
        break;
      case 4:
        D.set_vacuum_point(p.x + p.x0, p.y + p.y0);
        break;
      default:
        break;   // unknown point types are skipped with a warning in the reference
    }
  }
  int rc = D.fill(err, err_len);
  if (rc) return rc;
  D.trim();
  derive_flags(flags, nz, nr);
  return SPF_OK;
}

int spf_geometry_flags(int32_t nz, int32_t nr, int16_t *flags) {
  if (nz < 2 || nr < 2 || !flags) return SPF_EINVAL;
  derive_flags(flags, nz, nr);
  return SPF_OK;
}

}  // extern "C"
