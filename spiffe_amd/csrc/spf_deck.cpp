// spf_deck.cpp -- runs a spiffe input deck on the library: the command loop
// of src/spiffe.c:213-296 for the commands of the plain-method, gun and DC-gun
// configurations.  Host C++ over the C ABI only (every device action is an
// spf_* call); the deck text is read with the namelist reader, boundary,
// waveform, profile and saved-field files with the SDDS reader, and screen,
// lost-particle and field-saving files are written with the SDDS writer.
#include <float.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>
#include <strings.h>

#include <string>
#include <vector>

#include "../../include/spiffe_amd.h"
#include "spf_namelist.h"
#include "spf_sdds.h"

namespace spf {
namespace {

const double kPi = 3.141592653589793238462643;

struct DeckError {
  int code;
  std::string msg;
};

[[noreturn]] void bomb(const std::string &m, int code = SPF_EINVAL) { throw DeckError{code, m}; }

// compose_filename: "%s" in a file name stands for the run's rootname; a
// relative name is taken relative to the deck's directory
std::string compose(const std::string &name, const std::string &root, const std::string &dir) {
  std::string s = name;
  size_t p = s.find("%s");
  if (p != std::string::npos) s = s.substr(0, p) + root + s.substr(p + 2);
  if (!s.empty() && s[0] != '/' && !dir.empty()) s = dir + "/" + s;
  return s;
}

// match_string for the small keyword lists: case-insensitive unique prefix
int match(const std::string &v, const std::vector<const char *> &opts) {
  int hit = -1;
  for (size_t i = 0; i < opts.size(); i++)
    if (v.size() && strncasecmp(v.c_str(), opts[i], v.size()) == 0) {
      if (strlen(opts[i]) == v.size()) return (int)i;
      if (hit >= 0) return -1;
      hit = (int)i;
    }
  return hit;
}

void check_known(const Namelist &g, std::vector<const char *> known) {
  known.push_back(nullptr);
  std::string bad = g.unknown_keys(known.data());
  if (!bad.empty()) bomb("&" + g.group + ": unknown parameter(s) " + bad);
}

struct Antenna {
  int32_t direction = 0;
  int64_t istart = 0, iend = 0, iposition = 0;
  double current = 0, frequency = 0, phase = 0;
  std::vector<double> t, W;
  int64_t last_index = 0;
};

struct CathodeSpec {
  spf_cathode_desc d;
  std::vector<double> t, A;
};

struct ScreenSpec {
  std::string file;
  double z = 0;
  int32_t dir = 1;
  SddsWriter out;
};

struct Deck {
  std::string dir, root, name;
  spf_deck_options opt{};
  // geometry
  bool have_geometry = false;
  spf_run_desc d{};
  std::vector<int16_t> flags;
  std::vector<double> psi_imposed;
  bool tm = true, te = false;
  // fields from load_fields, [iz][ir]
  std::vector<double> F[12];
  bool have_field[12] = {};
  std::vector<Antenna> antennas;
  std::vector<CathodeSpec> cathodes;
  std::vector<spf_solenoid_desc> solenoids;
  std::vector<ScreenSpec *> screens;
  bool pc_set = false;
  spf_poisson_params pc{};
  // field saving
  std::string save_file;
  double save_interval = -1, save_start = 0, save_trigger = -1;
  int64_t save_before_exit = 0, save_step_interval = -1, save_start_step = -1;
  bool save_started = false;
  SddsWriter save_out;
  // run
  spf_ctx *ctx = nullptr;
  int32_t run = -1;
  int64_t integrations = 0;
  spf_deck_stats stats{};

  ~Deck() {
    for (auto *s : screens) delete s;
    if (ctx) {
      if (run >= 0) spf_run_destroy(ctx, run);
      spf_ctx_destroy(ctx);
    }
  }

  void say(const char *fmt, ...) {
    if (!opt.verbose) return;
    va_list ap;
    va_start(ap, fmt);
    vprintf(fmt, ap);
    va_end(ap);
    fflush(stdout);
  }

  void check(int rc) {
    if (rc) bomb(ctx ? spf_last_error(ctx) : "library call failed", rc);
  }

  static const char *field_name(int k) {
    static const char *n[12] = {"Ez", "Er", "Jz", "Jr", "Bphi", "Bphi2", "Ephi", "Jphi", "Br", "Br2", "Bz", "Bz2"};
    return n[k];
  }
  static int field_id(int k) {
    static const int id[12] = {SPF_EZ, SPF_ER, SPF_JZ, SPF_JR, SPF_BPHI, SPF_BPHI2,
                               SPF_EPHI, SPF_JPHI, SPF_BR, SPF_BR2, SPF_BZ, SPF_BZ2};
    return id[k];
  }

  // ---- define_geometry, src/geometry.c:21-369 ----
  void define_geometry(const Namelist &g) {
    check_known(g, {"nz", "nr", "zmin", "zmax", "rmax", "zr_factor", "rootname", "boundary", "boundary_output",
                    "urmel_boundary_output", "discrete_boundary_output", "interior_points", "lower", "upper",
                    "right", "left", "include_te_fields", "exclude_tm_fields", "turn_off_er", "turn_off_ez",
                    "turn_off_ephi", "turn_off_br", "turn_off_bz", "turn_off_bphi", "print_grids",
                    "radial_interpolation", "longitudinal_interpolation", "radial_smearing",
                    "longitudinal_smearing"});
    if (have_geometry) bomb("define_geometry given twice");
    std::string e, boundary, lower = "Dirichlet", upper = "Neumann", right = "Neumann", left = "Neumann";
    int64_t nz = 0, nr = 0, te = 0, xtm = 0, rad = 1, lon = 1, rs = 0, ls = 0;
    double zmin = 0, zmax = 0, rmax = 0, zr = 1;
    int64_t off[6] = {};
    g.get("nz", &nz, e);
    g.get("nr", &nr, e);
    g.get("zmin", &zmin, e);
    g.get("zmax", &zmax, e);
    g.get("rmax", &rmax, e);
    g.get("zr_factor", &zr, e);
    g.get("rootname", &root, e);
    g.get("boundary", &boundary, e);
    g.get("lower", &lower, e);
    g.get("upper", &upper, e);
    g.get("right", &right, e);
    g.get("left", &left, e);
    g.get("include_te_fields", &te, e);
    g.get("exclude_tm_fields", &xtm, e);
    g.get("radial_interpolation", &rad, e);
    g.get("longitudinal_interpolation", &lon, e);
    g.get("radial_smearing", &rs, e);
    g.get("longitudinal_smearing", &ls, e);
    const char *offn[6] = {"turn_off_ez", "turn_off_er", "turn_off_ephi", "turn_off_bz", "turn_off_br", "turn_off_bphi"};
    for (int k = 0; k < 6; k++) g.get(offn[k], &off[k], e);
    if (!e.empty()) bomb(e);
    if (boundary.empty()) bomb("no boundary-points file named");
    if (nz < 2 || nr < 2) bomb("nz and nr must be greater than 2");
    if (zmax <= zmin) bomb("zmax must be greater than zmin");
    if (rmax <= 0) bomb("rmax must be greater than 0");
    const std::vector<const char *> bc = {"Dirichlet", "Neumann"};
    int lo = match(lower, bc), up = match(upper, bc), ri = match(right, bc), le = match(left, bc);
    if (lo < 0) bomb("unrecognized boundary condition for lower boundary");
    if (lo == SPF_NEUMANN) bomb("Neumann boundary conditions are meaningless on the lower boundary");
    if (up < 0) bomb("unrecognized boundary condition for upper boundary");
    if (up == SPF_DIRICHLET) bomb("Dirichlet boundary conditions are not implemented for the upper boundary");
    if (ri < 0) bomb("unrecognized boundary condition for right boundary");
    if (le < 0) bomb("unrecognized boundary condition for left boundary");
    d.nz = (int32_t)nz;
    d.nr = (int32_t)nr;
    d.zmin = zmin;
    d.zmax = zmax;
    d.rmax = rmax;
    d.dz = (zmax - zmin) / (nz - 1);
    d.dr = rmax / (nr - 1);
    d.lower_bc = lo;
    d.upper_bc = up;
    d.right_bc = ri;
    d.left_bc = le;
    tm = !xtm;
    this->te = te != 0;
    // src/geometry.c:79
    d.modes = (rad ? SPF_RADIAL_INTERP : 0) | (lon ? SPF_LONGIT_INTERP : 0) | (tm ? SPF_TM_FIELDS : 0) |
              (te ? SPF_TE_FIELDS : 0) | (rs ? SPF_RADIAL_SMEAR : 0) | (ls ? SPF_LONGIT_SMEAR : 0);
    for (int k = 0; k < 6; k++) d.turn_off[k] = (int16_t)off[k];
    // the boundary file
    std::string path = compose(boundary, root, dir);
    FILE *fp = fopen(path.c_str(), "rb");
    if (!fp) bomb("cannot open boundary file " + path);
    std::string text;
    char buf[4096];
    size_t n;
    while ((n = fread(buf, 1, sizeof(buf), fp)) > 0) text.append(buf, n);
    fclose(fp);
    char err[512];
    int64_t npts = 0;
    if (spf_geometry_parse(text.c_str(), zr, nullptr, 0, &npts, err, sizeof(err))) bomb(err);
    std::vector<spf_boundary_point> pts(npts);
    if (spf_geometry_parse(text.c_str(), zr, pts.data(), npts, &npts, err, sizeof(err))) bomb(err);
    size_t cells = (size_t)(nz + 1) * (nr + 1);
    flags.assign(cells, 0);
    psi_imposed.assign(cells, 0.0);
    std::vector<int16_t> material(cells);
    if (spf_geometry_draw(d.nz, d.nr, zmin, zmax, rmax, pts.data(), npts, flags.data(), psi_imposed.data(),
                          material.data(), err, sizeof(err)))
      bomb(err);
    for (int k = 0; k < 12; k++) F[k].assign(cells, 0.0);
    have_geometry = true;
    say("geometry: %d x %d lines, %lld boundary points\n", d.nz, d.nr, (long long)npts);
  }

  void need_geometry(const char *what) {
    if (!have_geometry) bomb(std::string("can't process ") + what + " before define_geometry");
  }

  // ---- load_fields, src/load_fields.c:20-255 ----
  void load_fields(const Namelist &g) {
    need_geometry("load_fields");
    check_known(g, {"filename", "ez_peak", "factor", "time_threshold", "overlay"});
    std::string e, file;
    double Ez_peak = 0, factor = 1, time_threshold = 0;
    g.get("filename", &file, e);
    g.get("ez_peak", &Ez_peak, e);
    g.get("factor", &factor, e);
    g.get("time_threshold", &time_threshold, e);
    if (!e.empty()) bomb(e);
    if (file.empty()) bomb("no filename given for field samples");
    if (Ez_peak == 0 && factor == 0) bomb("Ez_peak and factor cannot both be zero");
    SddsFile f;
    std::string err;
    if (sdds_read(compose(file, root, dir), f, err)) bomb(err);
    const char *pars[] = {"time", "nr", "nz", "dr", "dz", "zmin", "zmax", "rmax", "TMfields", "TEfields"};
    for (const char *p : pars)
      if (f.param_index(p) < 0) bomb("load file is invalid--parameter " + std::string(p) + " missing");
    const SddsPage *pg = nullptr;
    for (const auto &p : f.pages)
      if (!(p.par_num[f.param_index("time")] < time_threshold)) {
        pg = &p;
        break;
      }
    if (!pg) bomb("unable to get data for field loading");
    int64_t nz0 = (int64_t)pg->par_num[f.param_index("nz")], nr0 = (int64_t)pg->par_num[f.param_index("nr")];
    double zmin0 = pg->par_num[f.param_index("zmin")], zmax0 = pg->par_num[f.param_index("zmax")];
    double rmax0 = pg->par_num[f.param_index("rmax")];
    bool ftm = pg->par_str[f.param_index("TMfields")] == "y", fte = pg->par_str[f.param_index("TEfields")] == "y";
    const int64_t nz1 = nz0 + 1, nr1 = nr0 + 1;
    if (pg->rows != nz1 * nr1) bomb("load file page has the wrong number of rows");
    std::vector<std::vector<double>> L(12);
    for (int k = 0; k < 12; k++) {
      bool want = k < 6 ? ftm : fte;
      if (!want) continue;
      int j = f.column_index(field_name(k));
      if (j < 0) {
        L[k].assign((size_t)(nz1 * nr1), 0.0);   // an absent column loads as zero
        continue;
      }
      L[k] = pg->cols[j].num;
    }
    if (Ez_peak) {
      if (!ftm) bomb("can't scale Ez_peak because TM mode fields are not enabled");
      double Ez_max = 0;
      for (int64_t iz = 1; iz <= nz0; iz++)
        if (fabs(Ez_max) < fabs(L[0][iz * nr1])) Ez_max = L[0][iz * nr1];
      if (fabs(Ez_max) < 1e-10) bomb("extremal value of on-axis Ez is < 1e-10--can't scale up");
      factor = Ez_peak / Ez_max;
    }
    if (factor != 1)
      for (int k = 0; k < 12; k++)
        for (auto &v : L[k]) v *= factor;
    // overlay onto the existing fields (always added, as in the reference)
    int64_t iz_start = 0, iz_end = d.nz, iz_offset = 0, ir_end = d.nr;
    double dz1 = (d.zmax - d.zmin) / (d.nz - 1), dz0 = (zmax0 - zmin0) / (nz0 - 1);
    double dr1 = d.rmax / (d.nr - 1), dr0 = rmax0 / (nr0 - 1);
    if (nz0 != d.nz || nr0 != d.nr || zmax0 != d.zmax || zmin0 != d.zmin || rmax0 != d.rmax) {
      if (fabs(dz0 / dz1 - 1) > 1e-6) bomb("dz's not close enough");
      if (fabs(dr0 / dr1 - 1) > 1e-6) bomb("dr's not close enough");
      iz_offset = (int64_t)((d.zmin - zmin0) / dz0 + 0.5);
      if (fabs(iz_offset - (d.zmin - zmin0) / dz0) > 1e-6)
        bomb("misalignment of boundary and grid lines from existing data and data in file");
      if (iz_offset < 0) iz_start = -iz_offset;
      if ((d.nz + iz_offset) > nz0) iz_end = nz0 - iz_offset;
      if (d.nr < nr0) ir_end = nr0;
      // the file's arrays bound what can be read
      if (ir_end > nr0) ir_end = nr0;
      if (ir_end > d.nr) ir_end = d.nr;
    }
    if (ftm && !tm) bomb("attempt to overlay TM mode fields onto problem which does not contain those fields");
    const int64_t nrh = d.nr + 1;
    for (int k = 0; k < 12; k++) {
      if (L[k].empty()) continue;
      if (k >= 6 && !te) continue;
      for (int64_t iz = iz_start; iz <= iz_end; iz++) {
        int64_t sz = iz + iz_offset;
        if (sz < 0 || sz > nz0 || iz > d.nz) continue;
        for (int64_t ir = 0; ir <= ir_end; ir++) F[k][iz * nrh + ir] += L[k][sz * nr1 + ir];
      }
      have_field[k] = true;
    }
    say("fields loaded from %s (factor %g)\n", file.c_str(), factor);
  }

  // ---- define_antenna, src/antenna.c:18-111 ----
  void define_antenna(const Namelist &g) {
    need_geometry("define_antenna");
    check_known(g, {"start", "end", "position", "direction", "current", "frequency", "phase", "waveform",
                    "time_offset"});
    std::string e, direction = "z", wf;
    double start = 0, end = 0, position = 0, time_offset = 0;
    Antenna a;
    g.get("start", &start, e);
    g.get("end", &end, e);
    g.get("position", &position, e);
    g.get("direction", &direction, e);
    g.get("current", &a.current, e);
    g.get("frequency", &a.frequency, e);
    g.get("phase", &a.phase, e);
    g.get("waveform", &wf, e);
    g.get("time_offset", &time_offset, e);
    if (!e.empty()) bomb(e);
    if (start >= end) bomb("start >= end for antenna ");
    if (direction.compare(0, 1, "r") == 0) {
      if (start < 0 || end > d.rmax) bomb("antenna start or end out of range");
      if (position < d.zmin || position > d.zmax) bomb("antenna position out of range");
      a.istart = (int64_t)(start / d.dr + 0.5);
      a.iend = (int64_t)(end / d.dr + 0.5);
      a.iposition = (int64_t)((position - d.zmin) / d.dz);
      a.direction = 1;
    } else if (direction.compare(0, 3, "phi") == 0) {
      if (start < d.zmin || end > d.zmax) bomb("antenna start or end out of range");
      if (position < 0 || position > d.rmax) bomb("antenna position out of range");
      a.istart = (int64_t)((start - d.zmin) / d.dz);
      a.iend = (int64_t)((end - d.zmin) / d.dz);
      a.iposition = (int64_t)(position / d.dr);
      a.direction = 2;
    } else if (direction.empty() || direction.compare(0, 1, "z") == 0) {
      if (start < d.zmin || end > d.zmax) bomb("antenna start or end out of range");
      if (position < 0 || position > d.rmax) bomb("antenna position out of range");
      a.istart = (int64_t)((start - d.zmin) / d.dz + 0.5);
      a.iend = (int64_t)((end - d.zmin) / d.dz + 0.5);
      a.iposition = (int64_t)(position / d.dr);
      a.direction = 0;
    } else {
      bomb("unknown antenna direction " + direction);
    }
    if (a.frequency <= 0) bomb("antenna frequency <= 0");
    if (wf.empty()) bomb("no antenna waveform file given");
    SddsFile f;
    std::string err;
    if (sdds_read(compose(wf, root, dir), f, err)) bomb(err);
    int it = f.column_index("t"), iw = f.column_index("W");
    if (f.pages.empty() || it < 0 || iw < 0 || f.pages[0].rows <= 2) bomb("SDDS waveform file empty or invalid");
    a.t = f.pages[0].cols[it].num;
    a.W = f.pages[0].cols[iw].num;
    for (auto &t : a.t) t += time_offset;
    for (size_t i = 1; i < a.t.size(); i++)
      if (a.t[i - 1] >= a.t[i]) bomb("waveform data is not ordered by increasing time index");
    antennas.push_back(a);
  }

  // antenna current at `time`, src/antenna.c:118-136
  double antenna_current(Antenna &a, double time) {
    const std::vector<double> &t = a.t, &W = a.W;
    int64_t n = (int64_t)t.size();
    double J;
    if (time <= t[0]) {
      J = W[0];
    } else if (time > t[n - 1] || a.last_index >= n - 1) {
      J = W[n - 1];
    } else {
      int64_t i = a.last_index;
      if (i != 0 && time < t[i]) bomb("antenna time indexing error");
      while (i < n - 2 && time >= t[i + 1]) i++;
      a.last_index = i;
      J = W[i] + (W[i + 1] - W[i]) / (t[i + 1] - t[i]) * (time - t[i]);
    }
    J *= a.current * sin(2 * kPi * a.frequency * time + a.phase);
    return J;
  }

  // ---- define_cathode (kept until the run exists; defined in deck order) ----
  void define_cathode(const Namelist &g) {
    need_geometry("define_cathode");
    check_known(g, {"z_position", "inner_radius", "outer_radius", "current_density", "temperature", "work_function",
                    "field_emission", "add_thermal_velocities", "field_emission_beta", "determine_temperature",
                    "electrons_per_macroparticle", "start_time", "stop_time", "autophase", "time_offset",
                    "number_per_step", "initial_pz", "initial_omega", "focal_length", "stiffness",
                    "discretize_radii", "random_number_seed", "distribution_correction_interval",
                    "spread_over_dt", "zoned_emission", "halton_radix_dt", "halton_radix_r", "profile",
                    "profile_time_name", "profile_factor_name", "emission_log"});
    CathodeSpec c;
    spf_cathode_defaults(&c.d);
    std::string e, profile, tname, fname;
    spf_cathode_desc &x = c.d;
    g.get("z_position", &x.z_position, e);
    g.get("inner_radius", &x.inner_radius, e);
    g.get("outer_radius", &x.outer_radius, e);
    g.get("current_density", &x.current_density, e);
    g.get("temperature", &x.temperature, e);
    g.get("work_function", &x.work_function, e);
    g.get("field_emission", &x.field_emission, e);
    g.get("add_thermal_velocities", &x.add_thermal_velocities, e);
    g.get("field_emission_beta", &x.field_emission_beta, e);
    g.get("determine_temperature", &x.determine_temperature, e);
    g.get("electrons_per_macroparticle", &x.electrons_per_macroparticle, e);
    g.get("start_time", &x.start_time, e);
    g.get("stop_time", &x.stop_time, e);
    g.get("autophase", &x.autophase, e);
    g.get("time_offset", &x.time_offset, e);
    g.get("number_per_step", &x.number_per_step, e);
    g.get("initial_pz", &x.initial_pz, e);
    g.get("initial_omega", &x.initial_omega, e);
    g.get("stiffness", &x.stiffness, e);
    g.get("discretize_radii", &x.discretize_radii, e);
    g.get("random_number_seed", &x.random_number_seed, e);
    g.get("distribution_correction_interval", &x.distribution_correction_interval, e);
    g.get("spread_over_dt", &x.spread_over_dt, e);
    g.get("zoned_emission", &x.zoned_emission, e);
    g.get("halton_radix_dt", &x.halton_radix_dt, e);
    g.get("halton_radix_r", &x.halton_radix_r, e);
    g.get("profile", &profile, e);
    g.get("profile_time_name", &tname, e);
    g.get("profile_factor_name", &fname, e);
    if (!e.empty()) bomb(e);
    if (!profile.empty()) {
      SddsFile f;
      std::string err;
      if (sdds_read(compose(profile, root, dir), f, err)) bomb(err);
      int it = f.column_index(tname.c_str()), ia = f.column_index(fname.c_str());
      if (f.pages.empty() || it < 0 || ia < 0) bomb("profile file lacks the named time or factor column");
      c.t = f.pages[0].cols[it].num;
      c.A = f.pages[0].cols[ia].num;
      if (c.t.empty()) bomb("no data in profile file");
    }
    x.e_mks = 1.602176634e-19;       // CODATA 2018, as the library's host mirror
    x.me_mks = 9.1093837015e-31;
    x.k_boltzmann_mks = 1.380649e-23;
    cathodes.push_back(c);
  }

  // ---- define_solenoid ----
  void define_solenoid(const Namelist &g) {
    need_geometry("define_solenoid");
    check_known(g, {"radius", "evaluation_radius_limit", "z_start", "z_end", "current", "bz_peak", "turns",
                    "symmetry", "field_output", "bucking", "z_buck", "gaussian", "gaussian_center",
                    "gaussian_sigma"});
    spf_solenoid_desc s;
    spf_solenoid_defaults(&s);
    std::string e;
    g.get("radius", &s.radius, e);
    g.get("evaluation_radius_limit", &s.evaluation_radius_limit, e);
    g.get("z_start", &s.z_start, e);
    g.get("z_end", &s.z_end, e);
    g.get("current", &s.current, e);
    g.get("bz_peak", &s.Bz_peak, e);
    g.get("turns", &s.turns, e);
    g.get("symmetry", &s.symmetry, e);
    g.get("bucking", &s.bucking, e);
    g.get("z_buck", &s.z_buck, e);
    g.get("gaussian", &s.gaussian, e);
    g.get("gaussian_center", &s.gaussian_center, e);
    g.get("gaussian_sigma", &s.gaussian_sigma, e);
    if (!e.empty()) bomb(e);
    solenoids.push_back(s);
  }

  // ---- define_screen, src/screens.c:29-117 ----
  void define_screen(const Namelist &g) {
    need_geometry("define_screen");
    check_known(g, {"filename", "template", "z_position", "delta_z", "number_of_screens", "start_time",
                    "direction"});
    std::string e, file, tmpl, direction = "forward";
    double z = 0, dz = 0, start_time = 0;
    int64_t n = 1;
    g.get("filename", &file, e);
    g.get("template", &tmpl, e);
    g.get("z_position", &z, e);
    g.get("delta_z", &dz, e);
    g.get("number_of_screens", &n, e);
    g.get("start_time", &start_time, e);
    g.get("direction", &direction, e);
    if (!e.empty()) bomb(e);
    if (n < 1) bomb("number_of_screens is invalid");
    if (file.empty() && n == 1) bomb("no filename given for screen");
    if (tmpl.empty() && n > 1) bomb("no template given for screens");
    if (!file.empty() && !tmpl.empty()) bomb("both filename and template are given--this doesn't make sense");
    int dc = match(direction, {"forward", "backward"});
    if (dc < 0) bomb("unknown screen direction specified");
    for (int64_t i = 0; i < n; i++) {
      if ((int)screens.size() >= 24) bomb("at most 24 screens");
      ScreenSpec *s = new ScreenSpec;
      if (!file.empty()) {
        s->file = compose(file, root, dir);
      } else {
        char nm[1024];
        snprintf(nm, sizeof(nm), compose(tmpl, root, dir).c_str(), (int)i);
        s->file = nm;
      }
      s->z = z;
      s->dir = dc == 0 ? 1 : -1;
      std::string err;
      std::vector<SddsField> pars = {{"direction", "", "", SDDS_T_STRING, dc == 0 ? "forward" : "backward"},
                                     {"Time", "s", "Time of output", SDDS_T_DOUBLE, ""},
                                     {"Timestep", "", "Time-step of output", SDDS_T_LONG, ""}};
      std::vector<SddsField> cols = {{"r", "m"}, {"pz", "m$be$nc"}, {"pr", "m$be$nc"}, {"pphi", "m$be$nc"},
                                     {"t", "s"}, {"q", "C"}, {"r0", "m"}, {"t0", "s"}};
      cols.push_back({"generation", "", "", SDDS_T_SHORT, ""});
      cols.push_back({"particleID", "", "", SDDS_T_LONG, ""});
      cols.push_back({"rMin", "m"});
      cols.push_back({"rMax", "m"});
      cols.push_back({"zMin", "m"});
      cols.push_back({"zMax", "m"});
      char label[512];
      snprintf(label, sizeof(label), "spiffe screen output for run %s for particles %s-crossing z=%e m",
               name.c_str(), dc == 0 ? "forward" : "backward", z);
      if (s->out.open(s->file, label, pars, cols, true, err)) bomb(err);
      screens.push_back(s);
      z += dz;
    }
  }

  // ---- poisson_correction, src/poisson_correction.c:31-72 ----
  void poisson_correction(const Namelist &g) {
    need_geometry("poisson_correction");
    check_known(g, {"start_time", "step_interval", "accuracy", "error_charge_threshold", "maximum_iterations",
                    "verbosity", "test_charge", "z_test_charge", "r_test_charge", "guess_type"});
    std::string e, guess = "none";
    double test_charge = 0;
    spf_poisson_params p{0, 0, 1e-6, 0, 1000, 0};
    int64_t verbosity = 0;
    g.get("start_time", &p.start_time, e);
    g.get("step_interval", &p.step_interval, e);
    g.get("accuracy", &p.accuracy, e);
    g.get("error_charge_threshold", &p.error_charge_threshold, e);
    g.get("maximum_iterations", &p.maximum_iterations, e);
    g.get("verbosity", &verbosity, e);
    g.get("test_charge", &test_charge, e);
    g.get("guess_type", &guess, e);
    if (!e.empty()) bomb(e);
    if (p.start_time < 0) bomb("start_time < 0 for poisson_correction");
    if (p.step_interval < 0) bomb("step_interval < 0 for poisson_correction");
    if (p.accuracy <= 0) bomb("accuracy <= 0 for poisson_correction");
    if (p.maximum_iterations <= 0) bomb("maximum iterations <= 0 for poisson_correction");
    if (test_charge) bomb("the test-charge check of poisson_correction is not supported", SPF_EUNSUPPORTED);
    int gc = match(guess, {"none", "line-charge", "point-charge", "zero"});
    if (gc < 0) bomb("unknown guess type for poisson_correction");
    p.guess_code = gc;
    if (!tm) bomb("TM fields must be enabled in order to do poisson correction");
    pc = p;
    pc_set = true;
  }

  // ---- define_field_saving, src/field_saving.c:17-43 ----
  void define_field_saving(const Namelist &g) {
    need_geometry("define_field_saving");
    check_known(g, {"filename", "time_interval", "start_time", "save_before_exiting", "z_ez_trigger"});
    std::string e;
    g.get("filename", &save_file, e);
    g.get("time_interval", &save_interval, e);
    g.get("start_time", &save_start, e);
    g.get("save_before_exiting", &save_before_exit, e);
    g.get("z_ez_trigger", &save_trigger, e);
    if (!e.empty()) bomb(e);
    if (save_file.empty()) bomb("no filename given for field samples");
    if (save_trigger != -1 && save_interval != -1) bomb("Either use time_interval or iz_Ez_trigger, not both");
    if (save_interval == -1 && save_trigger == -1) bomb("Use one of time_interval or iz_Ez_trigger");
    if (save_trigger != -1) bomb("z_Ez_trigger field saving is not supported", SPF_EUNSUPPORTED);
  }

  void set_constant_fields(const Namelist &g) {
    need_geometry("set_constant_fields");
    check_known(g, {"ez", "er", "ephi", "bz", "br", "bphi"});
    std::string e;
    const char *n[6] = {"ez", "er", "ephi", "bz", "br", "bphi"};
    for (int k = 0; k < 6; k++) {
      d.constant_fields[k] = 0;
      g.get(n[k], &d.constant_fields[k], e);
    }
    if (!e.empty()) bomb(e);
  }

  // field saving at the current step, src/field_saving.c:45-212
  void save_fields(int64_t step, double time, bool final_call) {
    if (save_file.empty()) return;
    if (!save_started) {
      save_step_interval = (int64_t)(save_interval / d.dt + 0.5);
      save_start_step = (int64_t)((save_start - d.start_time) / d.dt + 0.5);
      std::vector<SddsField> pars = {
          {"time", "s", "time of field save", SDDS_T_DOUBLE, ""},
          {"version", "", "save version", SDDS_T_LONG, "4"},
          {"nr", "", "number of r grid lines", SDDS_T_LONG, std::to_string(d.nr)},
          {"nz", "", "number of z grid lines", SDDS_T_LONG, std::to_string(d.nz)}};
      char b[64];
      auto num = [&](double v) {
        snprintf(b, sizeof(b), "%.17g", v);
        return std::string(b);
      };
      pars.push_back({"dr", "m", "spacing of r grid lines", SDDS_T_DOUBLE, num(d.dr)});
      pars.push_back({"dz", "m", "spacing of z grid lines", SDDS_T_DOUBLE, num(d.dz)});
      pars.push_back({"zmin", "m", "", SDDS_T_DOUBLE, num(d.zmin)});
      pars.push_back({"zmax", "m", "", SDDS_T_DOUBLE, num(d.zmax)});
      pars.push_back({"rmax", "m", "", SDDS_T_DOUBLE, num(d.rmax)});
      pars.push_back({"zt", "m", "", SDDS_T_DOUBLE, "0"});
      pars.push_back({"TMfields", "", "", SDDS_T_CHARACTER, tm ? "y" : "n"});
      pars.push_back({"TEfields", "", "", SDDS_T_CHARACTER, te ? "y" : "n"});
      std::vector<SddsField> cols = {{"metalFlags", "", "", SDDS_T_LONG, ""}};
      const char *units[12] = {"V/m", "V/m", "A/m$a2$n", "A/m$a2$n", "T", "T",
                               "V/m", "A/m$a2$n", "T", "T", "T", "T"};
      for (int k = 0; k < 12; k++)
        if ((k < 6 && tm) || (k >= 6 && te)) cols.push_back({field_name(k), units[k], "", SDDS_T_DOUBLE, ""});
      std::string err;
      if (save_out.open(compose(save_file, root, dir), "spiffe field saves", pars, cols, true, err)) bomb(err);
      save_started = true;
    }
    bool doSave = false;
    if (final_call && save_before_exit)
      doSave = true;
    else if (step >= save_start_step)
      doSave = save_step_interval == 0 || (step - save_start_step) % save_step_interval == 0;
    if (!doSave) return;
    const size_t cells = (size_t)(d.nz + 1) * (d.nr + 1);
    std::vector<std::vector<double>> a;
    std::vector<double> fl(flags.begin(), flags.end());
    std::vector<const void *> cols = {fl.data()};
    for (int k = 0; k < 12; k++)
      if ((k < 6 && tm) || (k >= 6 && te)) {
        a.emplace_back(cells);
        check(spf_field_download(ctx, run, field_id(k), a.back().data()));
      }
    for (auto &v : a) cols.push_back(v.data());
    char b[64];
    snprintf(b, sizeof(b), "%.17g", time);
    std::string err;
    if (save_out.write_page({b}, (int64_t)cells, cols, err)) bomb(err);
    stats.field_saves++;
  }

  // ---- integrate, src/integrate.c:24-208 ----
  void integrate(const Namelist &g) {
    need_geometry("integrate");
    check_known(g, {"dt_integration", "start_time", "finish_time", "status_interval", "space_charge",
                    "check_divergence", "smoothing_parameter", "j_filter_multiplier", "sg_smoothing_halfwidth",
                    "terminate_on_total_loss", "z_forces_start", "status_output", "lost_particles",
                    "imposed_efield_ramp_time", "imposed_efield_flat_top_time", "auto_max_dt"});
    std::string e, lost_file, status_file;
    double dt = 0, start_time = 0, finish_time = 0, smoothing = 0, jfm = 0, z_forces_start = -1;
    double ramp = 0, flat_top = 1;
    int64_t space_charge = 0, terminate = 0, sg = 0, auto_max_dt = 0;
    g.get("dt_integration", &dt, e);
    g.get("start_time", &start_time, e);
    g.get("finish_time", &finish_time, e);
    g.get("space_charge", &space_charge, e);
    g.get("smoothing_parameter", &smoothing, e);
    g.get("j_filter_multiplier", &jfm, e);
    g.get("sg_smoothing_halfwidth", &sg, e);
    g.get("terminate_on_total_loss", &terminate, e);
    g.get("z_forces_start", &z_forces_start, e);
    g.get("status_output", &status_file, e);
    g.get("lost_particles", &lost_file, e);
    g.get("imposed_efield_ramp_time", &ramp, e);
    g.get("imposed_efield_flat_top_time", &flat_top, e);
    g.get("auto_max_dt", &auto_max_dt, e);
    if (!e.empty()) bomb(e);
    if (dt <= 0) bomb("dt_integration <= 0");
    if (finish_time < start_time) bomb("finish_time < start_time");
    if (smoothing < 0 || smoothing > 1) bomb("smoothing_parameter must be >= 0 and < 1");
    if (jfm < 0 || jfm >= 1) bomb("J_filter_multiplier must be >=0 and < 1");
    if (smoothing || sg) bomb("current smoothing is not supported", SPF_EUNSUPPORTED);
    if (!status_file.empty()) say("note: status_output is not written\n");
    if (integrations++) bomb("a second integrate command is not supported", SPF_EUNSUPPORTED);
    if (space_charge) d.modes |= SPF_SPACE_CHARGE;
    d.start_time = start_time;
    d.dt = dt;
    int64_t n_steps = (int64_t)((finish_time - start_time) / dt);
    double dtC = (d.dz < d.dr ? d.dz : d.dr) / 299792458.0 / sqrt(2.0);
    if (dtC < d.dt) {
      if (!auto_max_dt) bomb("the time step must be less than the Courant limit for stable integration");
      d.dt = dtC;
      n_steps = (int64_t)((finish_time - start_time) / d.dt);
    }
    if (opt.max_steps > 0 && n_steps > opt.max_steps) n_steps = opt.max_steps;
    create_run();
    check(spf_set_z_forces_start(ctx, run, z_forces_start));
    check(spf_set_j_filter(ctx, run, jfm));
    // outputs
    SddsWriter lost_out;
    if (!lost_file.empty()) {
      std::vector<SddsField> cols = {{"t", "s"}, {"rLost", "m", "", SDDS_T_FLOAT, ""}, {"zLost", "m", "", SDDS_T_FLOAT, ""},
                                     {"pz", "", "", SDDS_T_FLOAT, ""}, {"pr", "", "", SDDS_T_FLOAT, ""},
                                     {"z0", "m", "", SDDS_T_FLOAT, ""}, {"r0", "m", "", SDDS_T_FLOAT, ""},
                                     {"q", "C", "", SDDS_T_FLOAT, ""}, {"generation", "", "", SDDS_T_SHORT, ""},
                                     {"particleID", "", "", SDDS_T_LONG, ""}};
      std::string err;
      if (lost_out.open(compose(lost_file, root, dir), "", {}, cols, true, err)) bomb(err);
    }
    // before the loop, src/integrate.c:125-133
    check(spf_set_step(ctx, run, 0));
    if (pc_set) {
      check(spf_imposed_efield(ctx, run, 1, ramp, flat_top));
      int64_t it;
      check(spf_poisson_correct(ctx, run, space_charge != 0, &it));
    }
    for (int32_t k = 0; k < n_cathodes; k++) check(spf_cathode_set_current_density(ctx, run, k));
    std::vector<int32_t> ldir(antennas.size());
    std::vector<int64_t> lpos(antennas.size()), lst(antennas.size()), len(antennas.size());
    std::vector<double> lJ(antennas.size());
    for (size_t a = 0; a < antennas.size(); a++) {
      ldir[a] = antennas[a].direction;
      lpos[a] = antennas[a].iposition;
      lst[a] = antennas[a].istart;
      len[a] = antennas[a].iend;
    }
    const int64_t drain_every = 16;
    bool seen = false;
    int64_t step = 0;
    for (; step < n_steps; step++) {
      double time = step * d.dt + d.start_time;
      check(spf_set_step(ctx, run, step));
      if (terminate) {
        int64_t np = 0;
        check(spf_particles_count(ctx, run, &np, nullptr));
        if (seen && !np) {
          say("simulation terminating due to total loss of particles.\n");
          break;
        }
        if (np) seen = true;
      }
      save_fields(step, time, false);
      for (size_t a = 0; a < antennas.size(); a++) lJ[a] = antennas[a].current ? antenna_current(antennas[a], time + d.dt / 2) : 0.0;
      check(spf_step(ctx, run, (int32_t)antennas.size(), ldir.data(), lpos.data(), lst.data(), len.data(), lJ.data()));
      if ((step + 1) % drain_every == 0) drain(lost_out);
    }
    drain(lost_out);
    save_fields(step, step * d.dt + d.start_time, true);
    stats.steps = step;
    check(spf_particles_count(ctx, run, &stats.particles, nullptr));
    int64_t next = 0;
    check(spf_particle_id_next(ctx, run, 0, &next));
    stats.emitted = next - 1;
  }

  int32_t n_cathodes = 0;

  void create_run() {
    int64_t cap = opt.max_particles;
    if (cap <= 0) {
      // room for everything the cathodes can emit: rate x window, plus 10%
      double total = 0;
      for (const auto &c : cathodes) {
        double area = kPi * (c.d.outer_radius * c.d.outer_radius - c.d.inner_radius * c.d.inner_radius);
        double n = c.d.number_per_step;
        if (c.d.electrons_per_macroparticle && area)
          n = area * d.dt * c.d.current_density / (c.d.e_mks * c.d.electrons_per_macroparticle);
        double steps = 1e9;
        if (c.d.stop_time > c.d.start_time) steps = (c.d.stop_time - c.d.start_time) / d.dt + 2;
        if (!c.t.empty()) steps = fmin(steps, (c.t.back() - c.t.front()) / d.dt + 2);
        total += (n * 1.2 + 2 * ((c.d.outer_radius - c.d.inner_radius) / d.dr + 2)) * steps;
      }
      if (total > 0x20000000) bomb("cannot size the particle arrays: give max_particles");
      cap = (int64_t)(total * 1.1) + 1024;
    }
    std::vector<double> sz;
    std::vector<int32_t> sd;
    for (auto *s : screens) {
      sz.push_back(s->z);
      sd.push_back(s->dir);
    }
    d.epsilon_o = 8.8541878128e-12;
    d.mu_o = 1.25663706212e-6;
    d.c_mks = 299792458.0;
    d.QoMC = -1.602176634e-19 / (9.1093837015e-31 * 299792458.0);
    d.stiffness = 1;
    d.metal_flags = flags.data();
    d.max_particles = cap;
    d.n_screens = (int32_t)screens.size();
    d.screen_z = sz.data();
    d.screen_direction = sd.data();
    d.compat = opt.compat;
    if (spf_ctx_create(opt.device, &ctx)) {
      ctx = nullptr;
      bomb("no usable GPU: spf_ctx_create failed", SPF_EHIP);
    }
    check(spf_run_create(ctx, &d, &run));
    for (int k = 0; k < 12; k++)
      if (have_field[k]) check(spf_field_upload(ctx, run, field_id(k), F[k].data()));
    check(spf_field_upload(ctx, run, SPF_PSI_IMPOSED, psi_imposed.data()));
    for (const auto &s : solenoids) check(spf_solenoid_add(ctx, run, &s, nullptr));
    for (auto &c : cathodes) {
      if (!c.t.empty()) {
        c.d.n_profile = (int64_t)c.t.size();
        c.d.profile_time = c.t.data();
        c.d.profile_factor = c.A.data();
      }
      int32_t first, count;
      check(spf_cathode_define(ctx, run, &c.d, &first, &count));
      n_cathodes = first + count;
    }
    if (pc_set) check(spf_poisson_setup(ctx, run, &pc));
    stats.max_particles = cap;
    say("run created: %lld particle slots, %d screens, %d cathodes, %zu antennas\n", (long long)cap,
        d.n_screens, n_cathodes, antennas.size());
  }

  // screen rows become one page per screen per step, lost rows one page per
  // step, in the order the push made them
  std::vector<spf_screen_row> rows;
  std::vector<spf_lost_row> lost;

  void drain(SddsWriter &lost_out) {
    // room for the device rings: every particle crossing every screen, every particle lost
    if (rows.empty()) rows.resize((size_t)((int64_t)screens.size() * stats.max_particles + 1024));
    if (lost.empty()) lost.resize((size_t)(stats.max_particles + 1024));
    int64_t nr_ = 0, nl_ = 0;
    check(spf_events_drain(ctx, run, rows.data(), (int64_t)rows.size(), &nr_, lost.data(), (int64_t)lost.size(), &nl_));
    stats.screen_rows += nr_;
    stats.lost_rows += nl_;
    std::string err;
    for (size_t s = 0; s < screens.size(); s++) {
      int64_t i = 0;
      while (i < nr_) {
        if (rows[i].screen != (int16_t)s) {
          i++;
          continue;
        }
        int64_t step = rows[i].step;
        std::vector<double> c[14];
        for (int64_t j = i; j < nr_ && rows[j].step == step; j++) {
          const spf_screen_row &q = rows[j];
          if (q.screen != (int16_t)s) continue;
          double v[14] = {q.r, q.pz, q.pr, q.pphi, q.t, q.q, q.r0, q.t0, (double)q.generation, (double)q.particleID,
                          q.rmin, q.rmax, q.zmin, q.zmax};
          for (int k = 0; k < 14; k++) c[k].push_back(v[k]);
        }
        int64_t j = i;
        while (j < nr_ && rows[j].step == step) j++;
        std::vector<const void *> cols;
        for (int k = 0; k < 14; k++) cols.push_back(c[k].data());
        char tb[64];
        snprintf(tb, sizeof(tb), "%.17g", step * d.dt + d.start_time);
        if (screens[s]->out.write_page({tb, std::to_string(step)}, (int64_t)c[0].size(), cols, err)) bomb(err);
        i = j;
      }
    }
    if (lost_out.is_open()) {
      int64_t i = 0;
      while (i < nl_) {
        int64_t step = lost[i].step, j = i;
        std::vector<double> c[10];
        for (; j < nl_ && lost[j].step == step; j++) {
          const spf_lost_row &q = lost[j];
          double v[10] = {q.t, q.rLost, q.zLost, q.pz, q.pr, q.z0, q.r0, q.q, (double)q.generation,
                          (double)q.particleID};
          for (int k = 0; k < 10; k++) c[k].push_back(v[k]);
        }
        std::vector<const void *> cols;
        for (int k = 0; k < 10; k++) cols.push_back(c[k].data());
        if (lost_out.write_page({}, j - i, cols, err)) bomb(err);
        i = j;
      }
    }
  }
};

}  // namespace
}  // namespace spf

using namespace spf;

extern "C" {

void spf_deck_defaults(spf_deck_options *o) {
  if (!o) return;
  memset(o, 0, sizeof(*o));
  o->compat = 1;
}

int spf_deck_run(const char *path, const spf_deck_options *opt, spf_deck_stats *stats, char *err, int64_t err_len) {
  auto report = [&](int code, const std::string &m) {
    if (err && err_len > 0) snprintf(err, (size_t)err_len, "%s", m.c_str());
    return code;
  };
  if (!path) return report(SPF_EINVAL, "no deck file");
  Deck D;
  if (opt) D.opt = *opt;
  else spf_deck_defaults(&D.opt);
  std::string p = path;
  size_t slash = p.rfind('/');
  D.dir = slash == std::string::npos ? "" : p.substr(0, slash);
  std::string base = slash == std::string::npos ? p : p.substr(slash + 1);
  D.name = base;
  D.root = base.substr(0, base.find('.'));
  FILE *fp = fopen(path, "rb");
  if (!fp) return report(SPF_EINVAL, "cannot open deck " + p);
  std::string text;
  char buf[4096];
  size_t n;
  while ((n = fread(buf, 1, sizeof(buf), fp)) > 0) text.append(buf, n);
  fclose(fp);
  std::vector<Namelist> groups;
  std::string e;
  if (parse_namelists(text, groups, e)) return report(SPF_EINVAL, p + ": " + e);
  try {
    for (const auto &g : groups) {
      const std::string &c = g.group;
      if (c == "define_geometry") D.define_geometry(g);
      else if (c == "load_fields") D.load_fields(g);
      else if (c == "define_antenna") D.define_antenna(g);
      else if (c == "define_cathode") D.define_cathode(g);
      else if (c == "define_solenoid") D.define_solenoid(g);
      else if (c == "define_screen") D.define_screen(g);
      else if (c == "poisson_correction") D.poisson_correction(g);
      else if (c == "define_field_saving") D.define_field_saving(g);
      else if (c == "set_constant_fields" || c == "constant_fields") D.set_constant_fields(g);
      else if (c == "integrate") D.integrate(g);
      else if (c == "stop") break;
      else if (c == "define_field_sampling" || c == "define_snapshots" || c == "define_field_output" ||
               c == "trace_execution")
        D.say("note: &%s is a diagnostic this runner does not write; skipped\n", c.c_str());
      else if (c == "define_secondary_emission" || c == "define_emitter" || c == "load_particles" ||
               c == "define_resistor" || c == "define_test_beam" || c == "add_on_axis_fields" ||
               c == "set_up_translation" || c == "space_charge")
        throw DeckError{SPF_EUNSUPPORTED, "&" + c + " is not supported by this runner"};
      else
        throw DeckError{SPF_EINVAL, "unknown namelist " + c};
    }
  } catch (const DeckError &x) {
    return report(x.code, x.msg);
  }
  if (stats) *stats = D.stats;
  return SPF_OK;
}

}  // extern "C"
