// Host file layer: input deck, PLOT3D, BC file, Overture file, and the
// reference's step loop with its report and solution-file cadence.
//
// Every reader follows the reference's own reader of that file, including
// its quirks (cited at each function); errors return LEE_EINVAL with a
// message instead of ending the process.

#include <sys/stat.h>

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "lee.h"
#include "lee_io.h"

namespace {

thread_local std::string g_io_error;

int io_fail(const char *fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_io_error = buf;
  return LEE_EINVAL;
}

// io::skip_this_line_of_textFile (src/io/io.cpp:520-560): empty lines,
// whitespace-only lines and lines whose first non-blank character is '#'.
bool skip_text_line(const std::string &s) {
  size_t i = 0;
  while (i < s.size() && isspace((unsigned char)s[i])) i++;
  return i == s.size() || s[i] == '#';
}

// The reference's line loop (io.cpp:583-593, 971-1001): getline, then
// `while (!eof)`; a last line that ends at end-of-file without a newline
// is never looked at.
bool read_text_lines(const std::string &path, std::vector<std::string> &out) {
  std::ifstream ifs(path.c_str(), std::ifstream::in);
  if (!ifs.is_open()) return false;
  std::string line;
  std::getline(ifs, line);
  while (!ifs.eof()) {
    if (!skip_text_line(line)) out.push_back(line);
    std::getline(ifs, line);
  }
  return true;
}

}  // namespace

struct lee_deck {
  struct Entry {
    std::string name;
    std::vector<std::string> body;
  };
  std::vector<Entry> e;

  const Entry *find(const char *name) const {
    for (const Entry &x : e)
      if (x.name == name) return &x;
    return nullptr;
  }
};

struct lee_overture {
  std::vector<int> recv_block, recv_l, donor_block, donor_l, widths, hole_block, hole_l;
  std::vector<double> coef;
  lee_overset view{};
};

extern "C" {

const char *lee_io_last_error(void) { return g_io_error.c_str(); }

// ------------------------------------------------------------------- deck

int lee_deck_open(const char *path, lee_deck **out) {
  if (!path || !out) return io_fail("bad argument");
  *out = nullptr;
  std::vector<std::string> lines;
  if (!read_text_lines(path, lines)) return io_fail("The input deck %s does not exist.", path);
  if (lines.empty()) return io_fail("The input deck %s has no entries.", path);
  // parse_linesInputDeck (src/core/input.cpp:870-999)
  // step 1: drop everything from the first '#'
  for (std::string &s : lines) {
    const size_t c = s.find('#');
    if (c != std::string::npos) s.erase(c);
  }
  // step 2: a backslash merges the line into the next one.  The reference
  // keeps substr(0, ibackslash - 1): the character before the backslash is
  // dropped too (and a backslash in the first column keeps the whole line).
  for (size_t l = 0; l < lines.size(); l++) {
    const size_t b = lines[l].find('\\');
    if (b == std::string::npos) continue;
    if (l + 1 == lines.size()) {
      lines[l].erase(b);
    } else {
      const std::string cur = lines[l].substr(0, (size_t)((long)b - 1));
      lines[l].clear();
      lines[l + 1] = cur + lines[l + 1];
    }
  }
  // steps 3 and 4: non-empty lines, split at the delimiters; a name seen
  // again blanks its earlier entries
  static const char *delim = " =,;:\t";
  lee_deck *d = new lee_deck;
  for (const std::string &s : lines) {
    if (s.empty()) continue;
    std::vector<std::string> tok;
    size_t p = 0;
    while (p < s.size()) {
      p = s.find_first_not_of(delim, p);
      if (p == std::string::npos) break;
      const size_t q = s.find_first_of(delim, p);
      tok.push_back(s.substr(p, q == std::string::npos ? std::string::npos : q - p));
      p = q;
    }
    if (tok.empty()) {
      delete d;
      return io_fail("An input-deck line holds only delimiters: \"%s\"", s.c_str());
    }
    for (lee_deck::Entry &x : d->e)
      if (x.name == tok[0]) {
        x.name.clear();
        x.body.clear();
      }
    d->e.push_back({tok[0], std::vector<std::string>(tok.begin() + 1, tok.end())});
  }
  *out = d;
  return LEE_OK;
}

void lee_deck_close(lee_deck *d) { delete d; }

int lee_deck_get(const lee_deck *d, const char *name, const char *keyword, int index, char *buf, int cap) {
  if (!d || !name || !buf || cap <= 0 || index < 0) return io_fail("bad argument");
  const lee_deck::Entry *x = d->find(name);
  if (!x) return io_fail("The name %s does not match any of the names in the input deck.", name);
  size_t at = 0;
  if (keyword) {
    size_t k = 0;
    while (k < x->body.size() && x->body[k] != keyword) k++;
    if (k == x->body.size())
      return io_fail("The keyword %s does not match any of the keywords in the list named as %s.", keyword, name);
    at = k + 1;
  }
  at += (size_t)index;
  if (at >= x->body.size()) return io_fail("The input-deck item %s has too few values.", name);
  const std::string &v = x->body[at];
  if ((int)v.size() >= cap) return io_fail("buffer too small for %s", name);
  memcpy(buf, v.c_str(), v.size() + 1);
  return LEE_OK;
}

}  // extern "C"

namespace {

struct DeckReader {
  const lee_deck *d;
  int rc = LEE_OK;
  std::string str(const char *name, const char *kw = nullptr) {
    char b[512];
    if (rc) return "";
    rc = lee_deck_get(d, name, kw, 0, b, sizeof b);
    return rc ? "" : std::string(b);
  }
  int i(const char *name, const char *kw = nullptr) { return atoi(str(name, kw).c_str()); }    // atoi as input.cpp:1026
  double f(const char *name, const char *kw = nullptr) { return atof(str(name, kw).c_str()); }  // atof as input.cpp:1115
};

bool copy_name(char *dst, const std::string &s) {
  if (s.size() >= 512) return false;
  memcpy(dst, s.c_str(), s.size() + 1);
  return true;
}

}  // namespace

extern "C" int lee_deck_settings(const lee_deck *d, lee_run_settings *s) {
  if (!d || !s) return io_fail("bad argument");
  memset(s, 0, sizeof *s);
  DeckReader r{d};
  lee_config &c = s->cfg;
  // UserInput::set_inputDeck (src/core/input.cpp:99-284), in its order
  c.num_dim = r.i("DIMENSION");
  const std::string axisym = r.str("AXISYMMETRY");
  const std::string model = r.str("PHYSICAL_MODEL");
  r.str("FLUID_MODEL");
  const std::string sim = r.str("SIMULATION");
  c.gamma = r.f("GAMMA_SPECIFICHEAT");
  const std::string fix = r.str("FIX_DT_OR_CFL");
  if (r.rc) return r.rc;
  std::string ax = axisym;  // truefalse_2int (input.cpp:786-800) compares in upper case
  for (char &ch : ax) ch = (char)toupper((unsigned char)ch);
  if (ax == "TRUE") {
    std::string axis = r.str("AXISYMMETRY", "AXIS_OF_SYMMETRY");  // xyz_2int (input.cpp:820-838)
    if (r.rc) return r.rc;
    for (char &ch : axis) ch = (char)toupper((unsigned char)ch);
    if (axis != "X") return io_fail("The axis of symmetry should be x by definition of the coordinate system.");
    if (c.num_dim != 2) return io_fail("Axisymmetry requires 2-D.");
    c.axisym = 1;
  } else if (ax != "FALSE") {
    return io_fail("AXISYMMETRY must be TRUE or FALSE.");
  }
  if (model == "LEE_SCALAR") {  // input.cpp:113-119, and the model/case check of :596-606, 620-621
    c.num_scalar = r.i("PHYSICAL_MODEL", "NUMBER_SCALAR");
    if (r.rc) return r.rc;
    if (!(c.num_scalar >= 1))
      return io_fail("For PHYSICAL_MODEL = LEE_SCALAR, at least one scalar should be solved for.");
    if (c.num_scalar > LEE_MAX_SCALAR) return io_fail("NUMBER_SCALAR = %d: the GPU advance carries at most %d.",
                                                      c.num_scalar, LEE_MAX_SCALAR);
    if (sim != "CASE_LINEAR_NOZZLE") return io_fail("Unknown simulation for the current physical model.");
  } else if (model != "LEE") {
    return io_fail("PHYSICAL_MODEL = %s is not supported by the GPU advance.", model.c_str());
  }
  if (!(c.gamma > 0.0)) return io_fail("GAMMA_SPECIFICHEAT must be positive.");
  if (fix == "FIX_DT") {
    c.fix_cfl = 0;
    c.cfl_or_dt = r.f("DT");
  } else if (fix == "FIX_CFL") {
    c.fix_cfl = 1;
    c.cfl_or_dt = r.f("CFL");
  } else {
    return io_fail("You may fix either DT or CFL in time integration.");
  }
  const std::string ts = r.str("TEMPORAL_SCHEME");
  s->num_time_steps = r.i("NUM_TOTAL_TIMESTEP");
  s->report_freq = r.i("REPORT_TIMESTEP");
  const std::string fd = r.str("SPATIAL_SCHEME");
  if (r.rc) return r.rc;
  if (ts == "RUNGE_KUTTA4") c.rk_stages = 4;
  else if (ts == "RUNGE_KUTTA2") c.rk_stages = 2;
  else return io_fail("TEMPORAL_SCHEME = %s is not supported.", ts.c_str());
  if (fd == "STANDARD_CENTRAL") c.fd_order = r.i("SPATIAL_SCHEME", "ORDER_ACCURACY");
  else if (fd == "FDO11P") c.fd_order = 11;
  else return io_fail("SPATIAL_SCHEME = %s is not supported.", fd.c_str());
  const std::string flt = r.str("SPATIAL_FILTER");
  if (r.rc) return r.rc;
  if (flt == "NONE") {
    c.filter_order = 0;
  } else if (flt == "STANDARD_CENTRAL") {
    c.filter_order = r.i("SPATIAL_FILTER", "ORDER_ACCURACY");
    c.filter_blend = r.f("SPATIAL_FILTER", "BLEND");
  } else if (flt == "SFO11P") {
    c.filter_order = 11;
    c.filter_strength = r.f("SPATIAL_FILTER", "STRENGTH");
    c.filter_blend = r.f("SPATIAL_FILTER", "BLEND");
  } else {
    return io_fail("The filter %s is not supported.", flt.c_str());
  }
  c.sponge_order = r.i("BUFFER_ZONE", "POLYNOMIAL_ORDER");
  c.sponge_constant = r.f("BUFFER_ZONE", "CONSTANT");
  r.str("METRIC_SCHEME");
  const std::string ov = r.str("OVERSET");
  if (r.rc) return r.rc;
  if (ov == "OVERTURE") {
    s->overset = 1;
    r.str("OVERSET", "TYPE");
    s->overset_order = r.i("OVERSET", "ORDER_ACCURACY");
  } else if (ov != "NONE") {
    return io_fail("%s is a unsupported format for overset-grid interpolation.", ov.c_str());
  }
  r.str("DOMAIN_DECOMPOSITION");
  const std::string fmt = r.str("FILE_FORMAT");
  const std::string grid = r.str("GRID_FILE");
  const std::string ovf = s->overset ? r.str("OVERSET_FILE") : std::string();
  const std::string mean = r.str("BASESTATE_FILE");
  const std::string bc = r.str("BC_FILE");
  s->time_writing_solutions = r.f("SOLUTION_WRITING_TIME");
  const std::string hw = r.str("HARMONIC_WAVE");
  const std::string interp = r.str("INTERPOLATE_SOLUTION");
  if (r.rc) return r.rc;
  if (fmt != "PLOT3D") return io_fail("Unknown type of data files.");
  if (hw != "NONE") {  // input.cpp:221-231, in its order
    const double amp = r.f("HARMONIC_WAVE", "AMPLITUDE");
    const std::string shape = r.str("HARMONIC_WAVE", "SHAPE");
    std::string dir = r.str("HARMONIC_WAVE", "DIRECTION");
    const double wl = r.f("HARMONIC_WAVE", "WAVELENGTH");
    double half = 0.0;
    if (shape == "WAVEFORM_GAUSSIAN_ROUNDJET") half = r.f("HARMONIC_WAVE", "HALFWIDTH");
    if (r.rc) return r.rc;
    for (char &ch : dir) ch = (char)toupper((unsigned char)ch);  // xyz_2int (input.cpp:820-838)
    if (dir != "X" && dir != "Y" && dir != "Z") return io_fail("Unknown direction %s.", dir.c_str());
    c.harmonic_wave = hw == "WAVE_ACOUSTIC" ? LEE_WAVE_ACOUSTIC : hw == "WAVE_ENTROPY" ? LEE_WAVE_ENTROPY : -1;
    if (c.harmonic_wave < 0) return io_fail("Unknown wave.");
    c.harmonic_shape = shape == "WAVEFORM_PLANE"               ? LEE_WAVEFORM_PLANE
                       : shape == "WAVEFORM_HOMOGENEOUS"       ? LEE_WAVEFORM_HOMOGENEOUS
                       : shape == "WAVEFORM_GAUSSIAN_ROUNDJET" ? LEE_WAVEFORM_GAUSSIAN_ROUNDJET
                                                               : 0;
    c.harmonic_direction = dir[0] - 'X';
    c.harmonic_amplitude = amp;
    c.harmonic_wavelength = wl;
    c.harmonic_halfwidth = half;
  }
  if (interp != "NONE") return io_fail("INTERPOLATE_SOLUTION = %s is not supported by the GPU advance.", interp.c_str());
  if (!copy_name(s->grid_file, grid) || !copy_name(s->mean_file, mean) || !copy_name(s->bc_file, bc) ||
      !copy_name(s->overset_file, ovf))
    return io_fail("file name too long");
  // source.cpp:22-107 acts only for these two cases
  c.source_kind = sim == "CASE_SCATTERING_TWOCYLINDER"         ? LEE_SOURCE_TWOCYLINDER
                  : sim == "CASE_2DJET_WITH_A_HARMONIC_SOURCE" ? LEE_SOURCE_JET2D
                                                               : LEE_SOURCE_NONE;
  c.precision = LEE_FP64;
  c.path = LEE_PATH_FUSED;
  if (s->report_freq <= 0) return io_fail("REPORT_TIMESTEP must be positive.");
  return LEE_OK;
}

// ----------------------------------------------------------------- PLOT3D

namespace {

struct P3D {
  int nblocks = 0;
  std::vector<int> dims;  // 4 per block
  long data_start = 0;
};

// A file kind and, for solution files, their variable count: LEE_P3D_SOLUTION
// holds 5, LEE_P3D_SOLUTION_VARS(nv) nv (false for an unknown kind).
bool p3d_kind(int &kind, int &nsol) {
  nsol = 5;
  if (kind >= LEE_P3D_SOLUTION_VARS(1) && kind <= LEE_P3D_SOLUTION_VARS(64)) {
    nsol = kind - LEE_P3D_SOLUTION_VARS(0);
    kind = LEE_P3D_SOLUTION;
  }
  return kind >= 0 && kind <= 2;
}

int p3d_header(FILE *f, int kind, int nsol, P3D &h) {
  int nb = 0;
  if (fread(&nb, 4, 1, f) != 1 || nb <= 0 || nb > (1 << 20)) return io_fail("bad PLOT3D header");
  h.nblocks = nb;
  h.dims.assign(4 * (size_t)nb, 0);
  const int per = kind == LEE_P3D_FUNCTION ? 4 : 3;
  for (int b = 0; b < nb; b++) {
    int v[4] = {0, 0, 0, kind == LEE_P3D_GRID ? 3 : nsol};
    if (fread(v, 4, per, f) != (size_t)per) return io_fail("truncated PLOT3D header");
    for (int q = 0; q < 4; q++) {
      if (v[q] <= 0) return io_fail("bad PLOT3D block size");
      h.dims[4 * b + q] = v[q];
    }
  }
  h.data_start = 4 + 4L * per * nb;
  return LEE_OK;
}

long p3d_block_bytes(int kind, const int *d) {
  const long n = (long)d[0] * d[1] * d[2];
  return 8L * d[3] * n + (kind == LEE_P3D_SOLUTION ? 32 : 0);
}

}  // namespace

extern "C" int lee_plot3d_dims(const char *path, int kind, int *nblocks, int *dims, int cap) {
  int nsol = 5;
  if (!path || !nblocks || !p3d_kind(kind, nsol)) return io_fail("bad argument");
  FILE *f = fopen(path, "rb");
  if (!f) return io_fail("The PLOT3D file %s does not exist.", path);
  P3D h;
  const int rc = p3d_header(f, kind, nsol, h);
  fclose(f);
  if (rc) return rc;
  *nblocks = h.nblocks;
  if (dims)
    for (int b = 0; b < std::min(cap, h.nblocks); b++)
      for (int q = 0; q < 4; q++) dims[4 * b + q] = h.dims[4 * b + q];
  return LEE_OK;
}

extern "C" int lee_plot3d_read(const char *path, int kind, int iblock, double *data, double *tau) {
  int nsol = 5;
  if (!path || !data || !p3d_kind(kind, nsol)) return io_fail("bad argument");
  FILE *f = fopen(path, "rb");
  if (!f) return io_fail("The PLOT3D file %s does not exist.", path);
  P3D h;
  int rc = p3d_header(f, kind, nsol, h);
  if (!rc && (iblock < 0 || iblock >= h.nblocks)) rc = io_fail("block %d is not in %s", iblock, path);
  if (!rc) {
    long off = h.data_start;
    for (int b = 0; b < iblock; b++) off += p3d_block_bytes(kind, &h.dims[4 * b]);
    const int *d = &h.dims[4 * iblock];
    const size_t cnt = (size_t)d[0] * d[1] * d[2] * d[3];
    double t4[4];
    if (fseek(f, off, SEEK_SET) != 0) rc = io_fail("seek failed in %s", path);
    else if (kind == LEE_P3D_SOLUTION && fread(t4, 8, 4, f) != 4) rc = io_fail("truncated PLOT3D file %s", path);
    else if (fread(data, 8, cnt, f) != cnt) rc = io_fail("truncated PLOT3D file %s", path);
    if (!rc && kind == LEE_P3D_SOLUTION && tau) memcpy(tau, t4, sizeof t4);
  }
  fclose(f);
  return rc;
}

extern "C" int lee_plot3d_write(const char *path, int kind, int nblocks, const int *dims, const double *const *data,
                                const double *tau) {
  int nsol = 5;
  if (!path || !dims || !data || nblocks <= 0 || !p3d_kind(kind, nsol)) return io_fail("bad argument");
  if (kind == LEE_P3D_SOLUTION && !tau) return io_fail("a solution file needs tau");
  for (int b = 0; b < nblocks; b++)
    if (dims[4 * b] <= 0 || dims[4 * b + 1] <= 0 || dims[4 * b + 2] <= 0 || dims[4 * b + 3] <= 0 || !data[b] ||
        (kind == LEE_P3D_GRID && dims[4 * b + 3] != 3) || (kind == LEE_P3D_SOLUTION && dims[4 * b + 3] != nsol))
      return io_fail("bad block %d", b);
  FILE *f = fopen(path, "wb");
  if (!f) return io_fail("cannot write %s", path);
  bool ok = fwrite(&nblocks, 4, 1, f) == 1;
  const int per = kind == LEE_P3D_FUNCTION ? 4 : 3;
  for (int b = 0; b < nblocks && ok; b++) ok = fwrite(&dims[4 * b], 4, per, f) == (size_t)per;
  for (int b = 0; b < nblocks && ok; b++) {
    const size_t cnt = (size_t)dims[4 * b] * dims[4 * b + 1] * dims[4 * b + 2] * dims[4 * b + 3];
    if (kind == LEE_P3D_SOLUTION) ok = fwrite(tau, 8, 4, f) == 4;
    if (ok) ok = fwrite(data[b], 8, cnt, f) == cnt;
  }
  if (fclose(f) != 0) ok = false;
  return ok ? LEE_OK : io_fail("write failed on %s", path);
}

// ---------------------------------------------------------------- BC file

namespace {

// io::parse_boundary_type (src/io/io.cpp:1294-1340), the types the advance
// implements
int bc_type(std::string s) {
  for (char &ch : s) ch = (char)tolower((unsigned char)ch);
  if (s == "dirichlet") return LEE_BC_DIRICHLET;
  if (s == "dirichlet_allzero") return LEE_BC_DIRICHLET_ALLZERO;
  if (s == "dirichlet_harmonicwave") return LEE_BC_DIRICHLET_HARMONICWAVE;
  if (s == "neumann") return LEE_BC_NEUMANN;
  if (s == "wall_slip_kinematic") return LEE_BC_WALL_SLIP_KINEMATIC;
  if (s == "centerline_cart_norm2x") return LEE_BC_CENTERLINE_CART_NORM2X;
  if (s == "centerline_cart_norm2y") return LEE_BC_CENTERLINE_CART_NORM2Y;
  if (s == "centerline_cart_norm2z") return LEE_BC_CENTERLINE_CART_NORM2Z;
  if (s == "centerline_axisym") return LEE_BC_CENTERLINE_AXISYM;
  if (s == "sponge_freund_ambient") return LEE_SPONGE_FREUND_AMBIENT;
  if (s == "sponge_freund_dirichlet") return LEE_SPONGE_FREUND_DIRICHLET;
  if (s == "sponge_freund_harmonicwave") return LEE_SPONGE_FREUND_HARMONICWAVE;
  return -1;
}

}  // namespace

extern "C" int lee_bc_file_read(const char *path, int nblocks, const int *dims, lee_boundary *out, int *iblock,
                                int cap, int *count) {
  if (!path || !dims || !count || nblocks <= 0 || cap < 0) return io_fail("bad argument");
  std::vector<std::string> lines;
  if (!read_text_lines(path, lines)) return io_fail("The file for your boundary treatment does not exist.");
  int nreg = 0;
  for (const std::string &s : lines) {
    // extract_bcData_out_of_ (io.cpp:1109-1270): entries separated by
    // blanks, exactly nine of them
    std::vector<std::string> tok;
    size_t p = 0;
    while (p < s.size() && tok.size() < 9) {
      p = s.find_first_not_of(' ', p);
      if (p == std::string::npos) break;
      const size_t q = s.find(' ', p);
      tok.push_back(s.substr(p, q == std::string::npos ? std::string::npos : q - p));
      p = q;
    }
    for (const std::string &t : tok)
      if (t[0] == '#') return io_fail("Some entry in the boundary condition file has a comment sign.");
    if (tok.size() != 9) return io_fail("Number of entries in the boundary condition file is incorrect.");
    const int blk = atoi(tok[0].c_str());
    if (blk < 0 || blk >= nblocks) return io_fail("The boundary condition file names block %d.", blk);
    lee_boundary r{};
    r.type = bc_type(tok[1]);
    if (r.type < 0) return io_fail("unsupported boundary type %s", tok[1].c_str());
    const std::string &sd = tok[2];
    if (sd[0] == '+') r.end = 0;
    else if (sd[0] == '-') r.end = 1;
    else return io_fail("Unknown side for a boundary in the boundary condition file.");
    std::string dn = sd.substr(1);
    for (char &ch : dn) ch = (char)tolower((unsigned char)ch);
    if (sd.size() == 3 && dn == "xi") r.dir = 0;
    else if (sd.size() == 4 && dn == "eta") r.dir = 1;
    else if (sd.size() == 5 && dn == "zeta") r.dir = 2;
    else return io_fail("Unknown direction for a boundary in the boundary condition file.");
    for (int e = 0; e < 6; e++) {
      const std::string &t = tok[3 + e];
      const int d = e / 2;
      int v = atoi((t[0] == '+' || t[0] == '-') ? t.c_str() + 1 : t.c_str());
      if (t[0] == '-') v = (dims[3 * blk + d] - 1) - v;
      (e % 2 == 0 ? r.is : r.ie)[d] = v;
    }
    if (nreg >= cap) return io_fail("more than %d regions in the boundary condition file", cap);
    if (out) out[nreg] = r;
    if (iblock) iblock[nreg] = blk;
    nreg++;
  }
  *count = nreg;
  return LEE_OK;
}

// --------------------------------------------------------------- Overture

extern "C" int lee_overture_open(const char *path, int num_dim, int nblocks, const int *dims, int order,
                                 lee_overture **out) {
  if (!path || !dims || !out || nblocks <= 0 || num_dim < 2 || num_dim > 3) return io_fail("bad argument");
  *out = nullptr;
  FILE *f = fopen(path, "rb");
  if (!f) return io_fail("The overset-grid file does not exist.");
  lee_overture *o = new lee_overture;
  auto fail = [&](const char *msg) {
    fclose(f);
    delete o;
    return io_fail("%s (%s)", msg, path);
  };
  auto rd = [&](void *p, size_t sz, size_t n) { return fread(p, sz, n, f) == n; };
  int nd = 0, nb = 0;
  if (!rd(&nd, 4, 1) || !rd(&nb, 4, 1)) return fail("truncated overset file");
  if (nd != num_dim) return fail("the overset file's dimension differs from the grid's");
  if (nb != nblocks) return fail("the overset file's block count differs from the grid's");
  std::vector<long> nn(nb);
  for (int b = 0; b < nb; b++) {
    int n = 0;
    if (!rd(&n, 4, 1)) return fail("truncated overset file");
    nn[b] = (long)dims[3 * b] * dims[3 * b + 1] * dims[3 * b + 2];
    if (n != nn[b]) return fail("the overset file's block size differs from the grid's");
  }
  std::vector<int> nrb(nb);
  long nrecv = 0;
  for (int b = 0; b < nb; b++) {
    if (!rd(&nrb[b], 4, 1) || nrb[b] < 0) return fail("bad receiver count");
    nrecv += nrb[b];
  }
  o->recv_block.resize(nrecv);
  if (nrecv && !rd(o->recv_block.data(), 4, nrecv)) return fail("truncated overset file");
  // receiver and donor (i, j[, k]) grouped by block, num_dim ints each
  auto read_ijk = [&](std::vector<int> &ijk) {
    ijk.assign(3 * nrecv, 0);
    for (long r = 0; r < nrecv; r++)
      if (!rd(&ijk[3 * r], 4, num_dim)) return false;
    return true;
  };
  std::vector<int> rijk, dijk;
  if (!read_ijk(rijk)) return fail("truncated overset file");
  o->donor_block.resize(nrecv);
  if (nrecv && !rd(o->donor_block.data(), 4, nrecv)) return fail("truncated overset file");
  if (!read_ijk(dijk)) return fail("truncated overset file");
  // per block: stencil widths, then that block's receivers' coefficients
  o->widths.assign(3 * nb, 1);
  std::vector<std::vector<double>> cb(nb);
  int stride = 1;
  for (int b = 0; b < nb; b++) {
    if (!rd(&o->widths[3 * b], 4, 3)) return fail("truncated overset file");
    for (int d = 0; d < num_dim; d++)
      if (o->widths[3 * b + d] - 1 != order) return fail("the stencil width does not match OVERSET ORDER_ACCURACY");
    const int sz = o->widths[3 * b] * o->widths[3 * b + 1] * o->widths[3 * b + 2];
    if (sz <= 0 || sz > 4096) return fail("bad stencil width");
    stride = std::max(stride, sz);
    cb[b].resize((size_t)nrb[b] * sz);
    if (!cb[b].empty() && !rd(cb[b].data(), 8, cb[b].size())) return fail("truncated overset file");
  }
  o->coef.assign((size_t)nrecv * stride, 0.0);
  long r0 = 0;
  for (int b = 0; b < nb; b++) {
    const int sz = o->widths[3 * b] * o->widths[3 * b + 1] * o->widths[3 * b + 2];
    for (int r = 0; r < nrb[b]; r++)
      std::copy(cb[b].begin() + (size_t)r * sz, cb[b].begin() + (size_t)(r + 1) * sz,
                o->coef.begin() + (size_t)(r0 + r) * stride);
    r0 += nrb[b];
  }
  std::vector<int> nh(nb);
  long nhole = 0;
  for (int b = 0; b < nb; b++) {
    if (!rd(&nh[b], 4, 1) || nh[b] < 0) return fail("bad hole count");
    nhole += nh[b];
  }
  for (int b = 0; b < nb; b++)
    for (int q = 0; q < nh[b]; q++) {
      int ijk[3];
      if (!rd(ijk, 4, 3)) return fail("truncated overset file");
      const int *d = &dims[3 * b];
      if (ijk[0] < 0 || ijk[0] >= d[0] || ijk[1] < 0 || ijk[1] >= d[1] || ijk[2] < 0 || ijk[2] >= d[2])
        return fail("hole outside its block");
      o->hole_block.push_back(b);
      o->hole_l.push_back(ijk[0] + d[0] * (ijk[1] + d[1] * ijk[2]));
    }
  fclose(f);
  // 1-D indices (geometry.h:131-136) in the receiver's and the donor's block
  o->recv_l.resize(nrecv);
  o->donor_l.resize(nrecv);
  for (long r = 0; r < nrecv; r++) {
    const int rb = o->recv_block[r], db = o->donor_block[r];
    if (rb < 0 || rb >= nb || db < 0 || db >= nb) {
      delete o;
      return io_fail("receiver %ld names a missing block (%s)", r, path);
    }
    const int *d = &dims[3 * rb], *e = &dims[3 * db];
    o->recv_l[r] = rijk[3 * r] + d[0] * (rijk[3 * r + 1] + d[1] * rijk[3 * r + 2]);
    o->donor_l[r] = dijk[3 * r] + e[0] * (dijk[3 * r + 1] + e[1] * dijk[3 * r + 2]);
  }
  lee_overset &v = o->view;
  v.nrecv = (int)nrecv;
  v.recv_block = o->recv_block.data();
  v.recv_l = o->recv_l.data();
  v.donor_block = o->donor_block.data();
  v.donor_l = o->donor_l.data();
  v.widths = o->widths.data();
  v.coef = o->coef.data();
  v.coef_stride = stride;
  v.nhole = (int)nhole;
  v.hole_block = o->hole_block.data();
  v.hole_l = o->hole_l.data();
  *out = o;
  return LEE_OK;
}

extern "C" const lee_overset *lee_overture_tables(const lee_overture *o) { return o ? &o->view : nullptr; }

extern "C" void lee_overture_close(lee_overture *o) { delete o; }

// -------------------------------------------------------------------- run

namespace {

std::string resolve(const std::string &dir, const char *name) {
  if (!name[0] || name[0] == '/' || dir.empty()) return name;
  return dir + "/" + name;
}

void mkdirs(const std::string &path) {
  std::string cur;
  for (size_t i = 0; i <= path.size(); i++) {
    if (i == path.size() || path[i] == '/') {
      if (!cur.empty()) mkdir(cur.c_str(), 0755);
    }
    if (i < path.size()) cur += path[i];
  }
}

struct Handle {
  lee_handle *h = nullptr;
  ~Handle() {
    if (h) lee_destroy(h);
  }
};

int lee_fail(lee_handle *h, int rc, const char *what) {
  g_io_error = std::string(what) + ": " + lee_last_error(h);
  return rc;
}

}  // namespace

extern "C" int lee_run_deck(const char *deck_path, const char *restart_path, int precision, int path,
                            const char *out_dir, int quiet) {
  if (!deck_path || !out_dir) return io_fail("bad argument");
  lee_deck *dk = nullptr;
  int rc = lee_deck_open(deck_path, &dk);
  if (rc) return rc;
  lee_run_settings st;
  rc = lee_deck_settings(dk, &st);
  lee_deck_close(dk);
  if (rc) return rc;
  st.cfg.precision = precision;
  st.cfg.path = path;
  std::string dir(deck_path);
  const size_t sl = dir.rfind('/');
  dir = sl == std::string::npos ? std::string() : dir.substr(0, sl);

  // variables: s', u', v', w', p' and the passive scalars (input.cpp:760-774)
  const int nv = 5 + st.cfg.num_scalar;
  // grid and base state (io::read_grid, io::read_function)
  const std::string gfile = resolve(dir, st.grid_file), mfile = resolve(dir, st.mean_file);
  int nb = 0;
  if ((rc = lee_plot3d_dims(gfile.c_str(), LEE_P3D_GRID, &nb, nullptr, 0))) return rc;
  std::vector<int> gd(4 * nb), md(4 * nb), d3(3 * nb);
  int nbm = 0;
  if ((rc = lee_plot3d_dims(gfile.c_str(), LEE_P3D_GRID, &nb, gd.data(), nb))) return rc;
  if ((rc = lee_plot3d_dims(mfile.c_str(), LEE_P3D_FUNCTION, &nbm, md.data(), nb))) return rc;
  if (nbm != nb) return io_fail("the base-state file has %d blocks, the grid %d", nbm, nb);
  std::vector<std::vector<double>> xyz(nb), mean(nb);
  for (int b = 0; b < nb; b++) {
    for (int q = 0; q < 3; q++) {
      if (md[4 * b + q] != gd[4 * b + q]) return io_fail("the base state and the grid differ in size (block %d)", b);
      d3[3 * b + q] = gd[4 * b + q];
    }
    // read_function_serialIO asserts num_vars_mean variables (plot3d.cpp:922)
    if (md[4 * b + 3] != nv) return io_fail("the base-state file must hold %d variables", nv);
    const size_t n = (size_t)gd[4 * b] * gd[4 * b + 1] * gd[4 * b + 2];
    xyz[b].resize(3 * n);
    mean[b].resize(nv * n);
    if ((rc = lee_plot3d_read(gfile.c_str(), LEE_P3D_GRID, b, xyz[b].data(), nullptr))) return rc;
    if ((rc = lee_plot3d_read(mfile.c_str(), LEE_P3D_FUNCTION, b, mean[b].data(), nullptr))) return rc;
  }
  // BC file
  std::vector<lee_boundary> regs(4096);
  std::vector<int> rblk(4096);
  int nreg = 0;
  if ((rc = lee_bc_file_read(resolve(dir, st.bc_file).c_str(), nb, d3.data(), regs.data(), rblk.data(), 4096, &nreg)))
    return rc;
  std::vector<std::vector<lee_boundary>> per(nb);
  for (int r = 0; r < nreg; r++) per[rblk[r]].push_back(regs[r]);
  std::vector<lee_block> blocks(nb);
  for (int b = 0; b < nb; b++)
    blocks[b] = {d3[3 * b], d3[3 * b + 1], d3[3 * b + 2], xyz[b].data(), mean[b].data(), (int)per[b].size(),
                 per[b].data()};
  // overset tables
  lee_overture *ovt = nullptr;
  if (st.overset &&
      (rc = lee_overture_open(resolve(dir, st.overset_file).c_str(), st.cfg.num_dim, nb, d3.data(), st.overset_order,
                              &ovt)))
    return rc;
  Handle H;
  rc = lee_create(&st.cfg, blocks.data(), nb, ovt ? lee_overture_tables(ovt) : nullptr, &H.h);
  lee_overture_close(ovt);
  if (rc) return lee_fail(nullptr, rc, "lee_create");

  // restart (io::read_solution): state and clock from tau; dt restarts at
  // the deck value (input.cpp:129-133)
  int step = 0;
  double t = 0.0;
  std::vector<std::vector<double>> sol(nb);
  for (int b = 0; b < nb; b++) sol[b].assign(nv * (size_t)d3[3 * b] * d3[3 * b + 1] * d3[3 * b + 2], 0.0);
  // io::read_solution reads num_vars_sol variables after tau (plot3d.cpp:715-833)
  const int rkind = nv == 5 ? LEE_P3D_SOLUTION : LEE_P3D_SOLUTION_VARS(nv);
  if (restart_path) {
    int nbr = 0;
    std::vector<int> rd(4 * nb);
    if ((rc = lee_plot3d_dims(restart_path, rkind, &nbr, rd.data(), nb))) return rc;
    if (nbr != nb) return io_fail("the restart file has %d blocks, the grid %d", nbr, nb);
    for (int b = 0; b < nb; b++) {
      for (int q = 0; q < 3; q++)
        if (rd[4 * b + q] != d3[3 * b + q]) return io_fail("the restart file and the grid differ in size (block %d)", b);
      double tau[4];
      if ((rc = lee_plot3d_read(restart_path, rkind, b, sol[b].data(), tau))) return rc;
      step = (int)tau[2];
      t = tau[3];
    }
  }
  for (int b = 0; b < nb; b++)
    if ((rc = lee_set_solution(H.h, b, sol[b].data()))) return lee_fail(H.h, rc, "lee_set_solution");
  if ((rc = lee_set_time(H.h, step, t, st.cfg.fix_cfl ? 1e-6 : st.cfg.cfl_or_dt)))
    return lee_fail(H.h, rc, "lee_set_time");
  const double t_lastrun = t;

  // the set-up writes of simulation.cpp:221-226: grid, initial solution
  // (counter 0) and base state with their name files (metrics and the
  // auxiliary variables are not written)
  const std::string od(out_dir);
  mkdirs(od + "/report/solution");
  // State::initialize_state_linearizedEuler (state.cpp:284-312): radial and
  // azimuthal names in an axisymmetric run
  const bool ax = st.cfg.axisym != 0;
  std::vector<std::string> names = {"S'", "UX'", ax ? "UR'" : "UY'", ax ? "UTHETA'" : "UZ'", "P'"};
  std::vector<std::string> mnames = {"Sbar", "UXbar", ax ? "URbar" : "UYbar", ax ? "UTHETAbar" : "UZbar", "Pbar"};
  {
    // State::initialize_state_linearizedEuler_scalar (state.cpp:345-357)
    // appends each counter to one stringstream: Z0', Z01', Z012', ...
    std::string counter;
    for (int k = 0; k < st.cfg.num_scalar; k++) {
      counter += std::to_string(k);
      names.push_back("Z" + counter + "'");
      mnames.push_back("Z" + counter + "bar");
    }
  }
  auto write_names = [&](const std::string &fn, const std::vector<std::string> &nm) {
    if (FILE *f = fopen(fn.c_str(), "w")) {
      for (const auto &v : nm) fprintf(f, "%s\n", v.c_str());
      fclose(f);
    }
  };
  auto write_fields = [&](const std::string &fn, int kind, int nvar, std::vector<std::vector<double>> &fields,
                          const double *tau) {
    std::vector<int> wd(4 * nb);
    std::vector<const double *> ptr(nb);
    for (int b = 0; b < nb; b++) {
      for (int q = 0; q < 3; q++) wd[4 * b + q] = d3[3 * b + q];
      wd[4 * b + 3] = nvar;
      ptr[b] = fields[b].data();
    }
    return lee_plot3d_write(fn.c_str(), kind, nb, wd.data(), ptr.data(), tau);
  };
  // io::write_solution (io.cpp:728-770): the PLOT3D solution format for 5
  // variables, the function format (no tau) for more
  auto write_solution = [&](const std::string &fn, const double *tau) {
    return nv == 5 ? write_fields(fn, LEE_P3D_SOLUTION, 5, sol, tau)
                   : write_fields(fn, LEE_P3D_FUNCTION, nv, sol, nullptr);
  };
  {
    const double tau0[4] = {0.0, 1.0, (double)step, t};
    if ((rc = write_fields(od + "/camille.xyz", LEE_P3D_GRID, 3, xyz, nullptr)) ||
        (rc = write_solution(od + "/camille.000000.q", tau0)) ||
        (rc = write_fields(od + "/mean_camille.q", LEE_P3D_FUNCTION, nv, mean, nullptr)))
      return rc;
    write_names(od + "/camille.nam", names);
    write_names(od + "/mean_camille.nam", mnames);
  }
  xyz.clear();
  mean.clear();

  // solver::solve with io::report_timeadvancing and write_solution_on_the_fly
  int counter = 0;
  while (step < st.num_time_steps) {
    const int next = std::min(st.num_time_steps, (step / st.report_freq + 1) * st.report_freq);
    if (next - step > 1 && (rc = lee_solve(H.h, next - step - 1))) return lee_fail(H.h, rc, "lee_solve");
    double dt_prev = 0.0, dt = 0.0;
    int s_now = 0;
    if ((rc = lee_get_time(H.h, &s_now, &t, &dt_prev))) return lee_fail(H.h, rc, "lee_get_time");
    if ((rc = lee_solve(H.h, 1)) || (rc = lee_synchronize(H.h))) return lee_fail(H.h, rc, "lee_solve");
    if ((rc = lee_get_time(H.h, &step, &t, &dt))) return lee_fail(H.h, rc, "lee_get_time");
    if (step % st.report_freq != 0) continue;
    // the step's CFL was evaluated with the previous dt (solver.cpp:53-64)
    double cfl = 0.0;
    if ((rc = lee_compute_cfl(H.h, &cfl))) return lee_fail(H.h, rc, "lee_compute_cfl");
    if (st.cfg.fix_cfl && dt > 0.0) cfl *= dt_prev / dt;
    if (!quiet) printf("Time step: %10d, time: %e, dt: %e, CFL: %e\n", step, t, dt, cfl);
    std::vector<double> lo(nv, INFINITY), hi(nv, -INFINITY);
    for (int b = 0; b < nb; b++) {
      if ((rc = lee_get_solution(H.h, b, sol[b].data()))) return lee_fail(H.h, rc, "lee_get_solution");
      const size_t n = sol[b].size() / nv;
      for (int v = 0; v < nv; v++)
        for (size_t p = 0; p < n; p++) {
          lo[v] = std::min(lo[v], sol[b][v * n + p]);
          hi[v] = std::max(hi[v], sol[b][v * n + p]);
        }
    }
    const std::string rep = od + "/report/solution/solution";
    if (FILE *f = fopen(rep.c_str(), "a")) {
      fprintf(f, "Time step: %10d, time: %e", step, t);
      for (int v = 0; v < nv; v++) fprintf(f, ", %s: %e ~ %e", names[v].c_str(), lo[v], hi[v]);
      fprintf(f, "\n");
      fclose(f);
    }
    if ((int)((t - t_lastrun) / st.time_writing_solutions) > counter || step == st.num_time_steps) {
      counter++;
      char fn[64];
      snprintf(fn, sizeof fn, "/camille.%06d.q", counter);
      const double tau[4] = {0.0, 1.0, (double)step, t};
      if ((rc = write_solution(od + fn, tau))) return rc;
      write_names(od + "/camille.nam", names);
    }
  }
  return LEE_OK;
}
