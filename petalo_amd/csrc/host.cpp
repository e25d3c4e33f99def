// host.cpp -- host-side parts of the C ABI: FOV helpers (src/fov.rs), the
// Image3D raw format (src/io/raw.rs:78-126), the MLEM config TOML
// (src/config/mlem.rs) and the seeded synthetic workloads.  No device code.
#include <hip/hip_runtime.h>

#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "../../include/petalo_hip.h"
#include "capi_internal.hpp"
#include "philox.hpp"
#include "siddon_walk.hpp"

using namespace petalo;

namespace {

struct ParseError : std::runtime_error {
  explicit ParseError(const std::string& m) : std::runtime_error(m) {}
};

template <class F> int guarded_host(F&& f) {
  try {
    return f();
  } catch (const ParseError& e) {
    return fail(PETALO_ERR_PARSE, e.what());
  } catch (const std::bad_alloc&) {
    return fail(PETALO_ERR_NOMEM, "host allocation failed");
  } catch (const std::exception& e) {
    return fail(PETALO_ERR_ARG, e.what());
  }
}

// ---- quantities ("200 ps", "3 cm"): uom FromStr = "<value> <unit>" ------
bool unit_factor(const std::string& u, char kind, double& f) {
  static const std::map<std::string, double> length = {
      {"mm", 1.0}, {"millimeter", 1.0}, {"millimeters", 1.0}, {"cm", 10.0}, {"centimeter", 10.0},
      {"centimeters", 10.0}, {"m", 1000.0}, {"meter", 1000.0}, {"meters", 1000.0}, {"km", 1e6},
      {"kilometer", 1e6}, {"kilometers", 1e6}, {"\xc2\xb5m", 1e-3}, {"um", 1e-3}, {"micrometer", 1e-3},
      {"micrometers", 1e-3}, {"nm", 1e-6}, {"nanometer", 1e-6}, {"nanometers", 1e-6}};
  static const std::map<std::string, double> time = {
      {"ps", 1.0}, {"picosecond", 1.0}, {"picoseconds", 1.0}, {"ns", 1e3}, {"nanosecond", 1e3},
      {"nanoseconds", 1e3}, {"\xc2\xb5s", 1e6}, {"us", 1e6}, {"microsecond", 1e6}, {"microseconds", 1e6},
      {"ms", 1e9}, {"millisecond", 1e9}, {"milliseconds", 1e9}, {"s", 1e12}, {"second", 1e12},
      {"seconds", 1e12}, {"fs", 1e-3}, {"femtosecond", 1e-3}, {"femtoseconds", 1e-3}};
  static const std::map<std::string, double> ratio = {{"", 1.0}, {"ratio", 1.0}, {"%", 0.01},
                                                      {"percent", 0.01}};
  const auto& m = kind == 'L' ? length : kind == 'T' ? time : ratio;
  auto it = m.find(u);
  if (it == m.end()) return false;
  f = it->second;
  return true;
}

std::string trim(const std::string& s) {
  size_t a = 0, b = s.size();
  while (a < b && std::isspace((unsigned char)s[a])) ++a;
  while (b > a && std::isspace((unsigned char)s[b - 1])) --b;
  return s.substr(a, b - a);
}

bool parse_f32(const std::string& s, float& v) {
  if (s.empty()) return false;
  char* end = nullptr;
  const double d = std::strtod(s.c_str(), &end);
  if (!end || *end != '\0') return false;
  v = (float)d;
  return true;
}

float quantity(const std::string& text, char kind) {
  const size_t sp = text.find(' ');
  if (sp == std::string::npos) {
    float v;
    if (kind == 'R' && parse_f32(text, v)) return v;
    throw ParseError("quantity '" + text + "' has no unit");
  }
  float v;
  if (!parse_f32(text.substr(0, sp), v)) throw ParseError("bad number in quantity '" + text + "'");
  double f;
  if (!unit_factor(trim(text.substr(sp + 1)), kind, f)) throw ParseError("unknown unit in quantity '" + text + "'");
  return (float)((double)v * f);
}

// ---- minimal TOML (the subset the mlem configs use) ----------------------
struct Value;
using Table = std::map<std::string, std::shared_ptr<Value>>;
struct Value {
  enum Kind { STR, INT, FLT, BOOL, ARR, TAB } kind = STR;
  std::string s;
  long long i = 0;
  double f = 0;
  bool b = false;
  std::vector<Value> a;
  Table t;
};

struct Toml {
  const std::string& src;
  size_t pos = 0;
  int line = 1;
  explicit Toml(const std::string& s) : src(s) {}

  [[noreturn]] void err(const std::string& m) {
    throw ParseError("TOML line " + std::to_string(line) + ": " + m);
  }
  bool eof() const { return pos >= src.size(); }
  char peek() const { return eof() ? '\0' : src[pos]; }
  void skip_ws() {  // spaces and tabs only
    while (!eof() && (src[pos] == ' ' || src[pos] == '\t')) ++pos;
  }
  void skip_comment() {
    if (peek() == '#')
      while (!eof() && src[pos] != '\n') ++pos;
  }
  void skip_ws_nl() {  // whitespace, newlines and comments
    for (;;) {
      skip_ws();
      skip_comment();
      if (peek() == '\n' || peek() == '\r') {
        if (peek() == '\n') ++line;
        ++pos;
        continue;
      }
      break;
    }
  }
  void end_of_line() {
    skip_ws();
    skip_comment();
    if (peek() == '\r') ++pos;
    if (!eof() && peek() != '\n') err("unexpected text after value");
  }
  std::string bare_key() {
    size_t a = pos;
    while (!eof() && (std::isalnum((unsigned char)peek()) || peek() == '_' || peek() == '-')) ++pos;
    if (a == pos) err("expected a key");
    return src.substr(a, pos - a);
  }
  std::string key_part() {
    if (peek() == '"') return string();
    return bare_key();
  }
  std::vector<std::string> dotted_key() {
    std::vector<std::string> k;
    skip_ws();
    k.push_back(key_part());
    for (;;) {
      skip_ws();
      if (peek() != '.') break;
      ++pos;
      skip_ws();
      k.push_back(key_part());
    }
    return k;
  }
  std::string string() {
    ++pos;  // opening quote
    std::string out;
    while (!eof() && peek() != '"') {
      char c = src[pos++];
      if (c == '\n') err("newline in string");
      if (c == '\\') {
        char e = src[pos++];
        switch (e) {
          case 'n': out += '\n'; break;
          case 't': out += '\t'; break;
          case '"': out += '"'; break;
          case '\\': out += '\\'; break;
          default: err("unsupported escape");
        }
      } else {
        out += c;
      }
    }
    if (eof()) err("unterminated string");
    ++pos;
    return out;
  }
  Value value() {
    skip_ws();
    Value v;
    const char c = peek();
    if (c == '"') {
      v.kind = Value::STR;
      v.s = string();
    } else if (c == '[') {
      v.kind = Value::ARR;
      ++pos;
      for (;;) {
        skip_ws_nl();
        if (peek() == ']') { ++pos; break; }
        v.a.push_back(value());
        skip_ws_nl();
        if (peek() == ',') { ++pos; continue; }
        if (peek() == ']') { ++pos; break; }
        err("expected ',' or ']' in array");
      }
    } else if (c == '{') {
      v.kind = Value::TAB;
      ++pos;
      skip_ws();
      if (peek() == '}') { ++pos; return v; }
      for (;;) {
        auto k = dotted_key();
        skip_ws();
        if (peek() != '=') err("expected '=' in inline table");
        ++pos;
        assign(v.t, k, value());
        skip_ws();
        if (peek() == ',') { ++pos; continue; }
        if (peek() == '}') { ++pos; break; }
        err("expected ',' or '}' in inline table");
      }
    } else if (src.compare(pos, 4, "true") == 0) {
      v.kind = Value::BOOL;
      v.b = true;
      pos += 4;
    } else if (src.compare(pos, 5, "false") == 0) {
      v.kind = Value::BOOL;
      pos += 5;
    } else {
      size_t a = pos;
      while (!eof() && (std::isalnum((unsigned char)peek()) || strchr("+-._", peek()))) ++pos;
      std::string tok = src.substr(a, pos - a), clean;
      if (tok.empty()) err("expected a value");
      for (char ch : tok)
        if (ch != '_') clean += ch;
      char* end = nullptr;
      const bool is_float = clean.find_first_of(".eE") != std::string::npos || clean == "inf" ||
                            clean == "+inf" || clean == "-inf" || clean == "nan";
      if (is_float) {
        v.kind = Value::FLT;
        v.f = std::strtod(clean.c_str(), &end);
      } else {
        v.kind = Value::INT;
        v.i = std::strtoll(clean.c_str(), &end, 10);
      }
      if (!end || *end != '\0') err("bad value '" + tok + "'");
    }
    return v;
  }
  void assign(Table& t, const std::vector<std::string>& key, Value v) {
    Table* cur = &t;
    for (size_t k = 0; k + 1 < key.size(); ++k) {
      auto& slot = (*cur)[key[k]];
      if (!slot) {
        slot = std::make_shared<Value>();
        slot->kind = Value::TAB;
      } else if (slot->kind != Value::TAB) {
        err("key '" + key[k] + "' is not a table");
      }
      cur = &slot->t;
    }
    auto& slot = (*cur)[key.back()];
    if (slot) err("duplicate key '" + key.back() + "'");
    slot = std::make_shared<Value>(std::move(v));
  }
  Table document() {
    Table root;
    Table* cur = &root;
    for (;;) {
      skip_ws_nl();
      if (eof()) break;
      if (peek() == '[') {
        ++pos;
        auto k = dotted_key();
        skip_ws();
        if (peek() != ']') err("expected ']'");
        ++pos;
        end_of_line();
        Table* t = &root;
        for (const auto& part : k) {
          auto& slot = (*t)[part];
          if (!slot) {
            slot = std::make_shared<Value>();
            slot->kind = Value::TAB;
          } else if (slot->kind != Value::TAB) {
            err("'" + part + "' is not a table");
          }
          t = &slot->t;
        }
        cur = t;
        continue;
      }
      auto k = dotted_key();
      skip_ws();
      if (peek() != '=') err("expected '='");
      ++pos;
      Value v = value();
      end_of_line();
      assign(*cur, k, std::move(v));
    }
    return root;
  }
};

const Value* get(const Table& t, const char* k) {
  auto it = t.find(k);
  return it == t.end() ? nullptr : it->second.get();
}

void deny_unknown(const Table& t, const std::vector<std::string>& allowed, const std::string& where) {
  for (const auto& kv : t) {
    bool ok = false;
    for (const auto& a : allowed) ok = ok || a == kv.first;
    if (!ok) throw ParseError("unknown field '" + kv.first + "' in " + where);
  }
}

const Table& as_table(const Value* v, const std::string& what) {
  if (!v || v->kind != Value::TAB) throw ParseError(what + " must be a table");
  return v->t;
}

std::string as_str(const Value* v, const std::string& what) {
  if (!v || v->kind != Value::STR) throw ParseError(what + " must be a string");
  return v->s;
}

double as_num(const Value* v, const std::string& what) {
  if (v && v->kind == Value::INT) return (double)v->i;
  if (v && v->kind == Value::FLT) return v->f;
  throw ParseError(what + " must be a number");
}

uint64_t as_uint(const Value* v, const std::string& what) {
  if (!v || v->kind != Value::INT || v->i < 0) throw ParseError(what + " must be a non-negative integer");
  return (uint64_t)v->i;
}

void copy_str(char* dst, size_t cap, const std::string& s, const std::string& what) {
  if (s.size() + 1 > cap) throw ParseError(what + " is too long");
  std::memcpy(dst, s.c_str(), s.size() + 1);
}

void bounds(const Table& in, const char* name, bool integer, int32_t& has_min, int32_t& has_max, double& mn,
            double& mx) {
  const Value* b = get(in, name);
  has_min = has_max = 0;
  if (!b) return;
  const Table& t = as_table(b, std::string("input.") + name);
  deny_unknown(t, {"min", "max"}, std::string("input.") + name);
  if (const Value* v = get(t, "min")) {
    mn = integer ? (double)as_uint(v, "min") : as_num(v, "min");
    has_min = 1;
  }
  if (const Value* v = get(t, "max")) {
    mx = integer ? (double)as_uint(v, "max") : as_num(v, "max");
    has_max = 1;
  }
}

// src/config/mlem.rs:73-253 (non-test build: missing mandatory fields fail).
void config_from_table(const Table& root, petalo_config_t& c) {
  std::memset(&c, 0, sizeof c);
  deny_unknown(root, {"input", "iterations", "fov", "tof", "attenuation_correction", "smear_energy",
                      "detector_full_axial_length", "scatter_correction"},
               "config");
  const Value* input = get(root, "input");
  if (!input) throw ParseError("missing mandatory field 'input'");
  {
    const Table& t = as_table(input, "input");
    deny_unknown(t, {"file", "dataset", "energy", "charge", "events"}, "input");
    if (!get(t, "file")) throw ParseError("missing mandatory field 'input.file'");
    if (!get(t, "dataset")) throw ParseError("missing mandatory field 'input.dataset'");
    copy_str(c.input_file, sizeof c.input_file, as_str(get(t, "file"), "input.file"), "input.file");
    copy_str(c.input_dataset, sizeof c.input_dataset, as_str(get(t, "dataset"), "input.dataset"), "input.dataset");
    double mn = 0, mx = 0;
    bounds(t, "energy", false, c.has_energy_min, c.has_energy_max, mn, mx);
    c.energy_min = (float)mn;
    c.energy_max = (float)mx;
    mn = mx = 0;
    bounds(t, "charge", false, c.has_charge_min, c.has_charge_max, mn, mx);
    c.charge_min = (float)mn;
    c.charge_max = (float)mx;
    mn = mx = 0;
    bounds(t, "events", true, c.has_events_min, c.has_events_max, mn, mx);
    c.events_min = (uint64_t)mn;
    c.events_max = (uint64_t)mx;
  }
  const Value* it = get(root, "iterations");
  if (!it) throw ParseError("missing mandatory field 'iterations'");
  {
    const Table& t = as_table(it, "iterations");
    deny_unknown(t, {"number", "subsets"}, "iterations");
    if (!get(t, "number")) throw ParseError("missing mandatory field 'iterations.number'");
    if (!get(t, "subsets")) throw ParseError("missing mandatory field 'iterations.subsets'");
    c.iterations_number = as_uint(get(t, "number"), "iterations.number");
    c.iterations_subsets = as_uint(get(t, "subsets"), "iterations.subsets");
  }
  const Value* fov = get(root, "fov");
  if (!fov) throw ParseError("missing mandatory field 'fov'");
  {
    const Table& t = as_table(fov, "fov");
    deny_unknown(t, {"nvoxels", "size"}, "fov");
    const Value* nv = get(t, "nvoxels");
    const Value* sz = get(t, "size");
    if (!nv) throw ParseError("missing mandatory field 'fov.nvoxels'");
    if (!sz) throw ParseError("missing mandatory field 'fov.size'");
    if (nv->kind != Value::ARR || nv->a.size() != 3) throw ParseError("fov.nvoxels must have 3 elements");
    if (sz->kind != Value::ARR || sz->a.size() != 3) throw ParseError("fov.size must have 3 elements");
    for (int i = 0; i < 3; ++i) {
      const uint64_t n = as_uint(&nv->a[i], "fov.nvoxels");
      if (n > 0xffffffffull) throw ParseError("fov.nvoxels too large");
      c.fov.n[i] = (uint32_t)n;
      c.fov.full_size_mm[i] = quantity(as_str(&sz->a[i], "fov.size"), 'L');
    }
  }
  if (const Value* tof = get(root, "tof")) {
    const Table& t = as_table(tof, "tof");
    deny_unknown(t, {"sigma", "cutoff"}, "tof");
    c.has_tof = 1;
    c.tof.enabled = 1;
    c.tof.sigma_ps = get(t, "sigma") ? quantity(as_str(get(t, "sigma"), "tof.sigma"), 'T') : 0.0f;
    c.tof.cutoff = get(t, "cutoff") ? (float)as_num(get(t, "cutoff"), "tof.cutoff") : 3.0f;  // mlem.rs:186
  }
  if (const Value* ac = get(root, "attenuation_correction")) {
    const Table& t = as_table(ac, "attenuation_correction");
    deny_unknown(t, {"sensitivity_image"}, "attenuation_correction");
    c.has_attenuation_correction = 1;
    if (const Value* p = get(t, "sensitivity_image"))
      copy_str(c.sensitivity_image, sizeof c.sensitivity_image, as_str(p, "sensitivity_image"), "sensitivity_image");
  }
  if (const Value* sc = get(root, "scatter_correction")) {
    const Table& t = as_table(sc, "scatter_correction");
    deny_unknown(t, {"phi", "r", "dz", "dt", "z"}, "scatter_correction");
    struct Spec {
      const char* name;
      std::vector<std::string> keys;
      char kind;
      const char* qty;
    };
    const Spec specs[] = {{"phi", {"bins"}, 0, nullptr},
                          {"r", {"bins", "max"}, 'L', "max"},
                          {"dz", {"bins", "max"}, 'L', "max"},
                          {"dt", {"bins", "max"}, 'T', "max"},
                          {"z", {"bins", "length"}, 'L', "length"}};
    for (const auto& s : specs) {
      const Value* v = get(t, s.name);
      if (!v) continue;
      const Table& st = as_table(v, std::string("scatter_correction.") + s.name);
      deny_unknown(st, s.keys, std::string("scatter_correction.") + s.name);
      for (const auto& k : s.keys)
        if (!get(st, k.c_str())) throw ParseError(std::string("missing field '") + k + "' in scatter_correction");
      const uint64_t bins = as_uint(get(st, "bins"), "bins");
      if (bins >= (1ull << 31)) throw ParseError("scatter_correction: too many bins");
      const float q = s.qty ? quantity(as_str(get(st, s.qty), s.qty), s.kind) : 0.0f;
      const std::string n = s.name;
      petalo_scattergram_spec_t& g = c.scatter;
      if (n == "phi") g.phi_bins = (uint32_t)bins;
      else if (n == "r") g.r_bins = (uint32_t)bins, g.r_max_mm = q;
      else if (n == "dz") g.dz_bins = (uint32_t)bins, g.dz_max_mm = q;
      else if (n == "dt") g.dt_bins = (uint32_t)bins, g.dt_max_ps = q;
      else g.z_bins = (uint32_t)bins, g.z_length_mm = q;
    }
    c.has_scatter_correction = 1;
  }
  if (const Value* se = get(root, "smear_energy")) {
    const Table& t = as_table(se, "smear_energy");
    deny_unknown(t, {"fwhm"}, "smear_energy");
    if (!get(t, "fwhm")) throw ParseError("missing field 'fwhm' in smear_energy");
    c.smear_energy_fwhm = quantity(as_str(get(t, "fwhm"), "smear_energy.fwhm"), 'R');
    c.has_smear_energy = 1;
  }
  if (const Value* dl = get(root, "detector_full_axial_length")) {
    const Table& t = as_table(dl, "detector_full_axial_length");
    deny_unknown(t, {"dz"}, "detector_full_axial_length");
    c.detector_full_axial_length_mm = quantity(as_str(get(t, "dz"), "detector_full_axial_length.dz"), 'L');
    c.has_detector_full_axial_length = 1;
  }
}

// ---- big-endian helpers --------------------------------------------------
void put_be16(std::string& s, uint16_t v) {
  s += (char)(v >> 8);
  s += (char)(v & 0xff);
}
void put_be32f(std::string& s, float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  s += (char)(u >> 24);
  s += (char)((u >> 16) & 0xff);
  s += (char)((u >> 8) & 0xff);
  s += (char)(u & 0xff);
}
uint16_t get_be16(const unsigned char* p) { return (uint16_t)((p[0] << 8) | p[1]); }
float get_be32f(const unsigned char* p) {
  const uint32_t u = ((uint32_t)p[0] << 24) | ((uint32_t)p[1] << 16) | ((uint32_t)p[2] << 8) | p[3];
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}

// ---- synthetic phantom ---------------------------------------------------
constexpr uint32_t kPhantomTag = 0x9e11u;
constexpr double kPi = 3.14159265358979323846;

int activity(double x, double y, double z) {
  static const double rad[6] = {5, 7, 9, 11, 13, 15};
  for (int k = 0; k < 6; ++k) {
    const double a = kPi / 3.0 * k;
    const double cx = 50.0 * std::cos(a), cy = 50.0 * std::sin(a);
    const double dx = x - cx, dy = y - cy;
    if (dx * dx + dy * dy + z * z <= rad[k] * rad[k]) return 4;
  }
  if (x * x + y * y <= 80.0 * 80.0 && std::fabs(z) <= 75.0) return 1;
  return 0;
}

void phantom_lor(uint64_t seed, uint32_t i, double sigma_dt, petalo_lor_t& out) {
  const double R = 350.0, HZ = 150.0, C = 0.299792458;
  for (uint32_t k = 0;; ++k) {
    const U32x4 a = philox(seed, i, k, 0, kPhantomTag);
    const U32x4 b = philox(seed, i, k, 1, kPhantomTag);
    const double r = 80.0 * std::sqrt(unit_f64(a.x));
    const double ph = 2.0 * kPi * unit_f64(a.y);
    const double x = r * std::cos(ph), y = r * std::sin(ph), z = 150.0 * (unit_f64(a.z) - 0.5);
    const int act = activity(x, y, z);
    if (!(unit_f64(a.w) * 4.0 < act)) continue;  // accept with probability activity/4
    const double ct = 2.0 * unit_f64(b.x) - 1.0;
    const double st = std::sqrt(std::max(0.0, 1.0 - ct * ct));
    const double pd = 2.0 * kPi * unit_f64(b.y);
    const double dx = st * std::cos(pd), dy = st * std::sin(pd), dz = ct;
    const double qa = dx * dx + dy * dy;
    if (qa < 1e-12) continue;
    const double qb = 2.0 * (x * dx + y * dy), qc = x * x + y * y - R * R;
    const double sq = std::sqrt(qb * qb - 4.0 * qa * qc);
    const double t1 = (-qb - sq) / (2.0 * qa), t2 = (-qb + sq) / (2.0 * qa);
    const double z1 = z + t1 * dz, z2 = z + t2 * dz;
    if (std::fabs(z1) > HZ || std::fabs(z2) > HZ) continue;
    double dt = (t2 - (-t1)) / C;  // (|e - p2| - |e - p1|) / c
    if (sigma_dt > 0) {
      const double u1 = ((double)b.z + 1.0) / 4294967296.0;  // (0, 1]
      const double u2 = unit_f64(b.w);
      dt += sigma_dt * std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * kPi * u2);
    }
    out.p1[0] = (float)(x + t1 * dx);
    out.p1[1] = (float)(y + t1 * dy);
    out.p1[2] = (float)z1;
    out.p2[0] = (float)(x + t2 * dx);
    out.p2[1] = (float)(y + t2 * dy);
    out.p2[2] = (float)z2;
    out.dt_ps = (float)dt;
    out.additive_correction = 1.0f;
    return;
  }
}

}  // namespace

extern "C" {

int petalo_fov_derive(const petalo_fov_t* fov, float half_width[3], float voxel_size[3]) {
  if (!fov || !half_width || !voxel_size) return fail(PETALO_ERR_ARG, "NULL argument");
  const DevFov f = make_fov(fov->full_size_mm, fov->n);
  half_width[0] = f.hw_x;
  half_width[1] = f.hw_y;
  half_width[2] = f.hw_z;
  voxel_size[0] = f.vs_x;
  voxel_size[1] = f.vs_y;
  voxel_size[2] = f.vs_z;
  return PETALO_OK;
}

int petalo_fov_voxel_centre(const petalo_fov_t* fov, const uint32_t i[3], float c[3]) {
  if (!fov || !i || !c) return fail(PETALO_ERR_ARG, "NULL argument");
  const DevFov f = make_fov(fov->full_size_mm, fov->n);
  c[0] = ((float)i[0] + 0.5f) * f.vs_x - f.hw_x;
  c[1] = ((float)i[1] + 0.5f) * f.vs_y - f.hw_y;
  c[2] = ((float)i[2] + 0.5f) * f.vs_z - f.hw_z;
  return PETALO_OK;
}

int petalo_fov_parse(const char* text, petalo_fov_t* out) {
  return guarded_host([&] {
    if (!text || !out) return fail(PETALO_ERR_ARG, "NULL argument");
    std::istringstream ss(text);
    std::vector<std::string> tok;
    std::string w;
    while (ss >> w) tok.push_back(w);
    if (tok.size() != 9) throw ParseError("FOV string needs 'dx unit dy unit dz unit nx ny nz'");
    for (int i = 0; i < 3; ++i) {
      out->full_size_mm[i] = quantity(tok[2 * i] + " " + tok[2 * i + 1], 'L');
      char* end = nullptr;
      const unsigned long long n = std::strtoull(tok[6 + i].c_str(), &end, 10);
      if (!end || *end || tok[6 + i][0] == '-' || n > 0xffffffffull) throw ParseError("bad voxel count");
      out->n[i] = (uint32_t)n;
    }
    return PETALO_OK;
  });
}

int petalo_parse_quantity(const char* text, char kind, float* value) {
  return guarded_host([&] {
    if (!text || !value) return fail(PETALO_ERR_ARG, "NULL argument");
    if (kind != 'L' && kind != 'T' && kind != 'R') return fail(PETALO_ERR_ARG, "kind must be 'L', 'T' or 'R'");
    *value = quantity(text, kind);
    return PETALO_OK;
  });
}

int petalo_config_parse(const char* text, petalo_config_t* out) {
  return guarded_host([&] {
    if (!text || !out) return fail(PETALO_ERR_ARG, "NULL argument");
    std::string s(text);
    Toml t(s);
    const Table root = t.document();
    config_from_table(root, *out);
    return PETALO_OK;
  });
}

int petalo_config_ingest(const petalo_config_t* c, float r_min_mm, float dr_mm, float dz_mm, float da_mm,
                         uint64_t seed, petalo_ingest_t* out) {
  if (!c || !out) return fail(PETALO_ERR_ARG, "NULL argument");
  // read_hdf5_lors unwraps config.smear_energy (hdf5.rs:107-108)
  if (!c->has_smear_energy) return fail(PETALO_ERR_ARG, "the config has no [smear_energy] table");
  petalo_ingest_t p;
  std::memset(&p, 0, sizeof p);
  p.has_events_min = c->has_events_min;
  p.has_events_max = c->has_events_max;
  p.events_min = c->events_min;
  p.events_max = c->events_max;
  p.has_charge_min = c->has_charge_min;
  p.has_charge_max = c->has_charge_max;
  p.charge_min = c->charge_min;
  p.charge_max = c->charge_max;
  p.has_energy_min = c->has_energy_min;
  p.has_energy_max = c->has_energy_max;
  p.energy_min = c->energy_min;
  p.energy_max = c->energy_max;
  p.smear_energy_fwhm = c->smear_energy_fwhm;
  // z_max = detector_full_axial_length.dz / 2 (hdf5.rs:99)
  p.has_z_max = c->has_detector_full_axial_length;
  p.z_max_mm = c->detector_full_axial_length_mm / 2.0f;
  // smear_positions: the dataset's discretization, Adjust::RandomZPhi (hdf5.rs:240-253)
  p.r_min_mm = r_min_mm;
  p.dr_mm = dr_mm;
  p.dz_mm = dz_mm;
  p.da_mm = da_mm;
  p.adjust = PETALO_ADJUST_RANDOM_Z_PHI;
  p.seed = seed;
  *out = p;
  return PETALO_OK;
}

int petalo_config_read(const char* path, petalo_config_t* out) {
  if (!path || !out) return fail(PETALO_ERR_ARG, "NULL argument");
  std::ifstream f(path, std::ios::binary);
  if (!f) return fail(PETALO_ERR_IO, std::string("cannot read config file ") + path);
  std::stringstream ss;
  ss << f.rdbuf();
  const std::string text = ss.str();
  return petalo_config_parse(text.c_str(), out);
}

int petalo_image3d_write(const char* path, const petalo_fov_t* fov, const float* data) {
  if (!path || !fov || !data) return fail(PETALO_ERR_ARG, "NULL argument");
  for (int i = 0; i < 3; ++i)
    if (fov->n[i] > 0xffff) return fail(PETALO_ERR_ARG, "Image3D stores voxel counts as u16");
  const size_t nv = (size_t)fov->n[0] * fov->n[1] * fov->n[2];
  std::string buf;
  buf.reserve(18 + 4 * nv);
  for (int i = 0; i < 3; ++i) put_be16(buf, (uint16_t)fov->n[i]);
  // From<&Image> for Image3D (raw.rs:105-115): mm = half_width * 2
  const DevFov f = make_fov(fov->full_size_mm, fov->n);
  put_be32f(buf, f.hw_x * 2.0f);
  put_be32f(buf, f.hw_y * 2.0f);
  put_be32f(buf, f.hw_z * 2.0f);
  for (size_t k = 0; k < nv; ++k) put_be32f(buf, data[k]);
  std::ofstream o(path, std::ios::binary | std::ios::trunc);
  if (!o) return fail(PETALO_ERR_IO, std::string("cannot write ") + path);
  o.write(buf.data(), (std::streamsize)buf.size());
  if (!o) return fail(PETALO_ERR_IO, std::string("write failed: ") + path);
  return PETALO_OK;
}

int petalo_image3d_read_header(const char* path, petalo_fov_t* fov_out) {
  if (!path || !fov_out) return fail(PETALO_ERR_ARG, "NULL argument");
  std::ifstream f(path, std::ios::binary);
  if (!f) return fail(PETALO_ERR_IO, std::string("cannot read ") + path);
  unsigned char h[18];
  f.read((char*)h, 18);
  if (f.gcount() != 18) return fail(PETALO_ERR_IO, "file too short for an Image3D header");
  for (int i = 0; i < 3; ++i) {
    fov_out->n[i] = get_be16(h + 2 * i);
    fov_out->full_size_mm[i] = get_be32f(h + 6 + 4 * i);
  }
  return PETALO_OK;
}

int petalo_image3d_read(const char* path, petalo_fov_t* fov_out, float* data, size_t capacity) {
  petalo_fov_t fov;
  int rc = petalo_image3d_read_header(path, &fov);
  if (rc) return rc;
  const size_t nv = (size_t)fov.n[0] * fov.n[1] * fov.n[2];
  if (fov_out) *fov_out = fov;
  if (!data) return PETALO_OK;
  if (capacity < nv) return fail(PETALO_ERR_SIZE, "capacity smaller than the image");
  std::ifstream f(path, std::ios::binary);
  f.seekg(18);
  std::vector<unsigned char> buf(4 * nv);
  f.read((char*)buf.data(), (std::streamsize)buf.size());
  if ((size_t)f.gcount() != buf.size()) return fail(PETALO_ERR_IO, "file shorter than its header says");
  for (size_t k = 0; k < nv; ++k) data[k] = get_be32f(buf.data() + 4 * k);
  return PETALO_OK;
}

int petalo_density_to_attenuation(float* image, size_t n, float rho_to_mu) {
  if (n && !image) return fail(PETALO_ERR_ARG, "image is NULL");
  // main.rs:11-17: the CLI value [cm^2/g] as an AreaPerMass in base units.
  const float g = 0.001f, cm = 10.0f;
  const float rho_unit = g / (cm * cm * cm);
  const float mu_unit = 1.0f / cm;
  const float apm = rho_to_mu * (mu_unit / rho_unit);
  // main.rs:80-94: divided by (1/mm) / (kg/m^3).
  const float kg = 1.0f, m = 1000.0f;
  const float rho_unit2 = kg / (m * m * m);
  const float mu_unit2 = 1.0f / 1.0f;
  const float factor = apm / (mu_unit2 / rho_unit2);
  for (size_t k = 0; k < n; ++k) image[k] *= factor;
  return PETALO_OK;
}

int petalo_synth_phantom_lors(uint64_t seed, size_t n, float fwhm_ps, petalo_lor_t* out) {
  if (n && !out) return fail(PETALO_ERR_ARG, "out is NULL");
  if (n >= (1ull << 32)) return fail(PETALO_ERR_ARG, "more than 2^32-1 LORs");
  if (!(fwhm_ps >= 0.0f)) return fail(PETALO_ERR_ARG, "fwhm must be >= 0");
  const double sigma_dt = (double)fwhm_ps / 2.3548200450309493;  // FWHM = 2 sqrt(2 ln 2) sigma
  const unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  std::vector<std::thread> th;
  for (unsigned t = 0; t < nt; ++t)
    th.emplace_back([=] {
      for (size_t i = t; i < n; i += nt) phantom_lor(seed, (uint32_t)i, sigma_dt, out[i]);
    });
  for (auto& x : th) x.join();
  return PETALO_OK;
}

int petalo_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  if (!ctr || !key || !out) return fail(PETALO_ERR_ARG, "NULL argument");
  const U32x4 r = philox4x32_10(U32x4{ctr[0], ctr[1], ctr[2], ctr[3]}, key[0], key[1]);
  out[0] = r.x;
  out[1] = r.y;
  out[2] = r.z;
  out[3] = r.w;
  return PETALO_OK;
}

int petalo_synth_density(const petalo_fov_t* fov, float* out) {
  if (!fov || !out) return fail(PETALO_ERR_ARG, "NULL argument");
  for (uint32_t z = 0; z < fov->n[2]; ++z)
    for (uint32_t y = 0; y < fov->n[1]; ++y)
      for (uint32_t x = 0; x < fov->n[0]; ++x) {
        const uint32_t i3[3] = {x, y, z};
        float c[3];
        petalo_fov_voxel_centre(fov, i3, c);
        float rho = 0.0f;
        const double cx = c[0], cy = c[1], cz = c[2];
        if (cx * cx + cy * cy <= 80.0 * 80.0 && std::fabs(cz) <= 75.0) rho = 1000.0f;
        if ((cx + 40.0) * (cx + 40.0) + cy * cy <= 15.0 * 15.0 && std::fabs(cz) <= 75.0) rho = 1791.67f;
        out[x + (y + (size_t)z * fov->n[1]) * fov->n[0]] = rho;
      }
  return PETALO_OK;
}

}  // extern "C"
