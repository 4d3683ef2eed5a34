// restart.hip -- the reference's restart record in UDF text: writer, reader,
// and the context save/load around them.  Host code only.
//
// Layout (UDF/define.udf:317-338): resume:{Calculation, CONTINUE:{Saved_Data:{
//   jikan:{ts, time}  Particles[]:sParticle  GR_body[]  GR_masses[]
//   GR_moments_body[]  Zeta[NX][NY][NZ_]:{zeta0, zeta1}  uk_dc:{x,y,z}
//   Concentration[N_spec][NX][NY][NZ_]:{ck}  oblique:{degree_oblique}}}}
// written by Save_Restart_udf / Save_Particle_udf (resume.cxx:10-78, 186-306)
// and read by Force_restore_parameters / Read_Particle_udf (resume.cxx:
// 308-562).  UDF data syntax (the reference's UDF/input.udf): records in
// braces, arrays in brackets, commas between adjacent scalars, strings in
// double quotes, // and /* */ comments.  Numbers are written with 17
// significant digits, so a save/load round trip is exact.
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "common.hpp"

namespace kapsel {

namespace {

// ------------------------------------------------------------------ writer
struct Out {
  std::string s;
  void num(double v) {
    char b[40];
    int n = snprintf(b, sizeof b, "%.17g", v);
    s.append(b, n);
    if (!strpbrk(b, ".eEn")) s += ".0";  // a float literal, as the schema types it
  }
  void integer(long long v) { s += std::to_string(v); }
  void vec(const double *v, int n) {
    s += '{';
    for (int i = 0; i < n; i++) {
      if (i) s += ',';
      num(v[i]);
    }
    s += '}';
  }
};

// sParticle member order (UDF/define.udf:283-306) and the state field each
// one comes from (-1: not kept by this library, written as zero)
constexpr int kSParticleVec[20] = {
    KAPSEL_P_X, KAPSEL_P_X_NOPBC, KAPSEL_P_V, KAPSEL_P_V_OLD, KAPSEL_P_F_HYDRO, KAPSEL_P_F_HYDRO_PREVIOUS,
    -1 /* f_hydro1 */, -1 /* f_slip */, -1 /* f_slip_previous */, KAPSEL_P_FR, KAPSEL_P_FR_PREVIOUS,
    KAPSEL_P_OMEGA, KAPSEL_P_OMEGA_OLD, KAPSEL_P_TORQUE_HYDRO, KAPSEL_P_TORQUE_HYDRO_PREVIOUS,
    -1 /* torque_hydro1 */, -1 /* torque_slip */, -1 /* torque_slip_previous */, KAPSEL_P_TORQUE_R,
    KAPSEL_P_TORQUE_R_PREVIOUS};

void write_resume(Out &o, const kapsel_restart_data &d) {
  const int np = d.n_particles, nz_ = 2 * (d.nz / 2 + 1);
  o.s += "resume:{\"CONTINUE\",\n {\n  {\n   {";
  o.integer(d.ts);
  o.s += ',';
  o.num(d.time);
  o.s += "}\n   [";
  for (int n = 0; n < np; n++) {
    o.s += "{";
    for (int f : kSParticleVec) {
      double v[3] = {0.0, 0.0, 0.0};
      if (f >= 0)
        for (int c = 0; c < 3; c++) v[c] = d.particles[((size_t)f * 3 + c) * np + n];
      o.vec(v, 3);
    }
    for (int h = 0; h < 2; h++) {
      double q[4];
      for (int c = 0; c < 4; c++) q[c] = d.orientations[(size_t)(4 * h + c) * np + n];
      o.vec(q, 4);
    }
    o.s += "}\n";
  }
  o.s += "]\n   []\n   []\n   []\n   [";
  for (int i = 0; i < d.nx; i++) {
    o.s += '[';
    for (int j = 0; j < d.ny; j++) {
      o.s += '[';
      for (int k = 0; k < nz_; k++) {
        const size_t im = ((size_t)i * d.ny + j) * nz_ + k;
        const double z[2] = {d.zeta0[im], d.zeta1[im]};
        o.vec(z, 2);
      }
      o.s += "]\n";
    }
    o.s += ']';
  }
  o.s += "]\n   ";
  o.vec(d.uk_dc, 3);
  o.s += "\n   [";
  for (int sp = 0; sp < d.n_spec; sp++) {
    const double *c = sp == 0 ? d.conc0 : d.conc1;
    o.s += '[';
    for (int i = 0; i < d.nx; i++) {
      o.s += '[';
      for (int j = 0; j < d.ny; j++) {
        o.s += '[';
        for (int k = 0; k < nz_; k++) o.vec(c + ((size_t)i * d.ny + j) * nz_ + k, 1);
        o.s += "]\n";
      }
      o.s += ']';
    }
    o.s += ']';
  }
  o.s += "]\n   {0.0}\n  }\n }\n}\n";
}

// ------------------------------------------------------------------ reader
struct In {
  const char *p, *end;
  std::string err;
  void fail(const char *what) {
    if (err.empty()) err = what;
    p = end;
  }
  bool ok() const { return err.empty(); }
  void skip() {  // whitespace, commas, comments
    while (p < end) {
      if (*p == ' ' || *p == '\t' || *p == '\r' || *p == '\n' || *p == ',') {
        p++;
      } else if (*p == '/' && p + 1 < end && p[1] == '/') {
        while (p < end && *p != '\n') p++;
      } else if (*p == '/' && p + 1 < end && p[1] == '*') {
        p += 2;
        while (p + 1 < end && !(p[0] == '*' && p[1] == '/')) p++;
        p = p + 2 <= end ? p + 2 : end;
      } else {
        break;
      }
    }
  }
  bool peek(char c) {
    skip();
    return p < end && *p == c;
  }
  void expect(char c) {
    skip();
    if (p < end && *p == c) p++;
    else fail(c == '{' ? "expected '{'" : c == '}' ? "expected '}'" : c == '[' ? "expected '['" : "expected ']'");
  }
  std::string str() {
    skip();
    if (p >= end || *p != '"') {
      fail("expected a string");
      return "";
    }
    const char *b = ++p;
    while (p < end && *p != '"') p++;
    std::string r(b, p);
    if (p < end) p++;
    return r;
  }
  double num() {
    skip();
    errno = 0;
    char *e = nullptr;
    std::string tok;
    while (p < end && !strchr(" \t\r\n,{}[]:\"", *p)) tok += *p++;
    double v = strtod(tok.c_str(), &e);
    if (tok.empty() || *e != '\0' || errno == ERANGE) fail("expected a number");
    return v;
  }
  // skips one value of any kind (records, arrays, strings, scalars)
  void value() {
    skip();
    if (p >= end) return fail("unexpected end of file");
    if (*p == '{' || *p == '[') {
      const char close = *p == '{' ? '}' : ']';
      p++;
      while (ok() && !peek(close)) value();
      expect(close);
    } else if (*p == '"') {
      str();
    } else {
      while (p < end && !strchr(" \t\r\n,{}[]:\"", *p)) p++;
    }
  }
  void vec(double *v, int n) {
    expect('{');
    for (int i = 0; i < n; i++) v[i] = num();
    expect('}');
  }
};

// the value range of the top-level record `resume` in the data section
bool find_resume(const std::string &text, size_t &b, size_t &e) {
  size_t start = text.find("\\begin{data}");
  start = start == std::string::npos ? 0 : start + strlen("\\begin{data}");
  In in{text.data() + start, text.data() + text.size(), ""};
  while (in.ok()) {
    in.skip();
    if (in.p >= in.end || *in.p == '\\') return false;  // \end{data}
    const char *kb = in.p;
    while (in.p < in.end && !strchr(" \t\r\n:{}[]", *in.p)) in.p++;
    std::string key(kb, in.p);
    in.skip();
    if (in.p >= in.end || *in.p != ':') return false;
    in.p++;
    in.skip();
    const char *vb = in.p;
    in.value();
    if (!in.ok()) return false;
    if (key == "resume") {
      b = vb - text.data();
      e = in.p - text.data();
      return true;
    }
  }
  return false;
}

void read_resume(In &in, kapsel_restart_data &d) {
  const int np = d.n_particles, nz_ = 2 * (d.nz / 2 + 1);
  in.expect('{');
  const std::string calc = in.str();
  if (!in.ok()) return;
  if (calc != "CONTINUE") return in.fail("resume.Calculation is not CONTINUE: the file holds no saved state");
  in.expect('{');  // CONTINUE
  in.expect('{');  // Saved_Data
  in.expect('{');  // jikan
  d.ts = (int)in.num();
  d.time = in.num();
  in.expect('}');
  in.expect('[');  // Particles[]
  int n = 0;
  while (in.ok() && !in.peek(']')) {
    if (n >= np) return in.fail("more saved particles than the context holds");
    in.expect('{');
    for (int f : kSParticleVec) {
      double v[3];
      in.vec(v, 3);
      if (f >= 0)
        for (int c = 0; c < 3; c++) d.particles[((size_t)f * 3 + c) * np + n] = v[c];
    }
    for (int c = 0; c < 3; c++)  // not saved: x_previous = x (resume.cxx:409-410)
      d.particles[((size_t)KAPSEL_P_X_PREVIOUS * 3 + c) * np + n] = d.particles[((size_t)KAPSEL_P_X * 3 + c) * np + n];
    for (int h = 0; h < 2; h++) {
      double q[4];
      in.vec(q, 4);
      const double qn = sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
      if (!(fabs(qn - 1.0) < 1e-6)) return in.fail("a saved orientation is not a unit quaternion");
      for (int c = 0; c < 4; c++) d.orientations[(size_t)(4 * h + c) * np + n] = q[c] / qn;  // qtn_normalize
    }
    in.expect('}');
    n++;
  }
  in.expect(']');
  if (in.ok() && n != np) return in.fail("the file holds a different number of particles");
  for (int r = 0; r < 3; r++) in.value();  // GR_body[], GR_masses[], GR_moments_body[]
  in.expect('[');                          // Zeta[NX][NY][NZ_]
  for (int i = 0; in.ok() && i < d.nx; i++) {
    if (in.peek(']')) return in.fail("Zeta: fewer planes than NX of the context");
    in.expect('[');
    for (int j = 0; in.ok() && j < d.ny; j++) {
      if (in.peek(']')) return in.fail("Zeta: fewer rows than NY of the context");
      in.expect('[');
      for (int k = 0; in.ok() && k < nz_; k++) {
        if (in.peek(']')) return in.fail("Zeta: rows shorter than NZ_ of the context");
        double z[2];
        in.vec(z, 2);
        const size_t im = ((size_t)i * d.ny + j) * nz_ + k;
        d.zeta0[im] = z[0];
        d.zeta1[im] = z[1];
      }
      if (in.ok() && !in.peek(']')) return in.fail("Zeta: rows longer than NZ_ of the context");
      in.expect(']');
    }
    if (in.ok() && !in.peek(']')) return in.fail("Zeta: more rows than NY of the context");
    in.expect(']');
  }
  if (in.ok() && !in.peek(']')) return in.fail("Zeta: more planes than NX of the context");
  in.expect(']');
  in.vec(d.uk_dc, 3);
  in.expect('[');  // Concentration[N_spec][NX][NY][NZ_]
  int ns = 0;
  while (in.ok() && !in.peek(']')) {
    if (ns >= 2) return in.fail("Concentration: more than two species");
    double *c = ns == 0 ? d.conc0 : d.conc1;
    const bool keep = ns < d.n_spec && c;  // read only for an electrolyte (resume.cxx:327-333)
    in.expect('[');
    for (int i = 0; in.ok() && i < d.nx; i++) {
      in.expect('[');
      for (int j = 0; in.ok() && j < d.ny; j++) {
        in.expect('[');
        for (int k = 0; in.ok() && k < nz_; k++) {
          double v;
          in.vec(&v, 1);
          if (keep) c[((size_t)i * d.ny + j) * nz_ + k] = v;
        }
        in.expect(']');
      }
      in.expect(']');
    }
    in.expect(']');
    ns++;
  }
  in.expect(']');
  if (in.ok() && d.n_spec > 0 && ns < d.n_spec) return in.fail("the file holds no ion concentrations");
  in.value();  // oblique
  in.expect('}');
  in.expect('}');
  in.expect('}');
}

bool read_file(const char *path, std::string &text) {
  FILE *f = fopen(path, "rb");
  if (!f) return false;
  char buf[1 << 16];
  size_t n;
  while ((n = fread(buf, 1, sizeof buf, f)) > 0) text.append(buf, n);
  fclose(f);
  return true;
}

void check_data(const kapsel_restart_data *d) {
  if (!d || d->nx < 1 || d->ny < 1 || d->nz < 2 || d->n_particles < 0 || (d->n_spec != 0 && d->n_spec != 2) ||
      !d->zeta0 || !d->zeta1 || (d->n_particles > 0 && (!d->particles || !d->orientations)) ||
      (d->n_spec == 2 && (!d->conc0 || !d->conc1))) {
    set_error("invalid restart data (sizes or null buffers)");
    throw Error{KAPSEL_ERR_ARG};
  }
}

}  // namespace

void restart_write(const char *path, const char *template_path, const kapsel_restart_data *d) {
  check_data(d);
  if (!path) {
    set_error("null path");
    throw Error{KAPSEL_ERR_ARG};
  }
  Out o;
  write_resume(o, *d);
  std::string text;
  if (template_path) {
    if (!read_file(template_path, text)) {
      set_error("cannot read the template %s", template_path);
      throw Error{KAPSEL_ERR_ARG};
    }
    size_t b, e;
    if (find_resume(text, b, e)) {
      // drop the template's "resume:" key too: the record text carries it
      size_t kb = text.rfind("resume", b);
      text = text.substr(0, kb) + o.s + text.substr(e);
    } else {
      size_t end = text.find("\\end{data}");
      if (end == std::string::npos) text += o.s;
      else text.insert(end, o.s);
    }
  } else {
    text = "\\include{\"define.udf\"}\n\\begin{data}\n" + o.s + "\\end{data}\n";
  }
  FILE *f = fopen(path, "wb");
  if (!f || fwrite(text.data(), 1, text.size(), f) != text.size()) {
    if (f) fclose(f);
    set_error("cannot write %s", path);
    throw Error{KAPSEL_ERR_ARG};
  }
  if (fclose(f) != 0) {
    set_error("cannot write %s", path);
    throw Error{KAPSEL_ERR_ARG};
  }
}

void restart_read(const char *path, kapsel_restart_data *d) {
  check_data(d);
  std::string text;
  if (!path || !read_file(path, text)) {
    set_error("cannot read %s", path ? path : "(null)");
    throw Error{KAPSEL_ERR_ARG};
  }
  size_t b, e;
  if (!find_resume(text, b, e)) {
    set_error("%s: no resume record in the data section", path);
    throw Error{KAPSEL_ERR_ARG};
  }
  In in{text.data() + b, text.data() + e, ""};
  read_resume(in, *d);
  if (!in.ok()) {
    set_error("%s: %s (at byte %zu)", path, in.err.c_str(), (size_t)(in.p - text.data()));
    throw Error{KAPSEL_ERR_ARG};
  }
}

}  // namespace kapsel
