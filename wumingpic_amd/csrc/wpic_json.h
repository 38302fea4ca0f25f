// wpic_json.h -- a small ordered JSON tree with JSON-Fortran's print layout
// and a reader, for the raw + JSON files (wpic_io.cpp) and the app driver's
// config (wpic_app.cpp).  Host code only.
#pragma once

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <utility>
#include <vector>

namespace wpic {
namespace json {

// ---------------------------------------------------------------------------
// JSON values (objects keep insertion order, as JSON-Fortran's do)
// ---------------------------------------------------------------------------
struct JVal {
  enum Kind { OBJ, ARR, INT, REAL, STR, BOOL, NUL } k = NUL;
  long long i = 0;
  double d = 0.0;
  bool b = false;
  std::string s;
  std::vector<std::pair<std::string, JVal>> obj;
  std::vector<JVal> arr;

  static JVal object() {
    JVal v;
    v.k = OBJ;
    return v;
  }
  static JVal integer(long long x) {
    JVal v;
    v.k = INT;
    v.i = x;
    return v;
  }
  static JVal real(double x) {
    JVal v;
    v.k = REAL;
    v.d = x;
    return v;
  }
  static JVal string(const std::string& x) {
    JVal v;
    v.k = STR;
    v.s = x;
    return v;
  }
  JVal& add(const std::string& key, JVal v) {
    obj.emplace_back(key, std::move(v));
    return obj.back().second;
  }
  const JVal* get(const std::string& key) const {
    if (k != OBJ) return nullptr;
    for (auto& kv : obj)
      if (kv.first == key) return &kv.second;
    return nullptr;
  }
};

// A real as JSON-Fortran writes it by default: write(str,'(ss,E27.17E4)')
// (json_parameters.F90:76 with the precision of json_initialize) followed by
// compact_real_string (json_string_utilities.F90:266-334), which drops
// trailing zeros of the significand (keeping one after the point) and
// leading zeros of the exponent (keeping one).  E editing gives the
// significand in [0.1, 1): 17 correctly rounded significant digits.
// Non-finite values are written quoted, as json%initialize leaves
// non_normals_to_null false.
inline std::string fmt_real(double v) {
  if (std::isnan(v)) return "\"NaN\"";
  if (std::isinf(v)) return v > 0 ? "\"Infinity\"" : "\"-Infinity\"";
  char buf[64];
  std::snprintf(buf, sizeof buf, "%.16e", v);
  const char* p = buf;
  const bool neg = *p == '-';
  if (neg) p++;
  std::string digits(1, *p++);
  if (*p == '.') p++;
  while (*p && *p != 'e') digits += *p++;
  const int e10 = *p == 'e' ? std::atoi(p + 1) : 0;
  const int fexp = v == 0.0 ? 0 : e10 + 1;
  const size_t last = digits.find_last_not_of('0');
  const std::string sig = last == std::string::npos ? std::string("0") : digits.substr(0, last + 1);
  std::string out = neg ? "-0." : "0.";
  out += sig;
  out += fexp < 0 ? "E-" : "E+";
  out += std::to_string(fexp < 0 ? -fexp : fexp);
  return out;
}

inline std::string quote(const std::string& s) {
  std::string o = "\"";
  for (char ch : s) {
    switch (ch) {
      case '"': o += "\\\""; break;
      case '\\': o += "\\\\"; break;
      case '\n': o += "\\n"; break;
      case '\r': o += "\\r"; break;
      case '\t': o += "\\t"; break;
      case '\b': o += "\\b"; break;
      case '\f': o += "\\f"; break;
      default:
        if ((unsigned char)ch < 0x20) {
          char u[8];
          std::snprintf(u, sizeof u, "\\u%04x", (unsigned)(unsigned char)ch);
          o += u;
        } else {
          o += ch;
        }
    }
  }
  return o + "\"";
}

inline void print(const JVal& v, int indent, std::string& out) {
  const std::string pad(indent, ' '), pad2(indent + 2, ' ');
  switch (v.k) {
    case JVal::OBJ:
      if (v.obj.empty()) {
        out += "{}";
        return;
      }
      out += "{\n";
      for (size_t n = 0; n < v.obj.size(); n++) {
        out += pad2 + quote(v.obj[n].first) + ": ";
        print(v.obj[n].second, indent + 2, out);
        out += n + 1 < v.obj.size() ? ",\n" : "\n";
      }
      out += pad + "}";
      return;
    case JVal::ARR:
      if (v.arr.empty()) {
        out += "[]";
        return;
      }
      out += "[\n";
      for (size_t n = 0; n < v.arr.size(); n++) {
        out += pad2;
        print(v.arr[n], indent + 2, out);
        out += n + 1 < v.arr.size() ? ",\n" : "\n";
      }
      out += pad + "]";
      return;
    case JVal::INT: out += std::to_string(v.i); return;
    case JVal::REAL: out += fmt_real(v.d); return;
    case JVal::STR: out += quote(v.s); return;
    case JVal::BOOL: out += v.b ? "true" : "false"; return;
    default: out += "null"; return;
  }
}

// ---------------------------------------------------------------------------
// JSON reading (for paraio__input): objects, arrays, numbers, strings
// ---------------------------------------------------------------------------
struct Parser {
  const char* p;
  const char* e;
  bool ok = true;
  void ws() {
    while (p < e && (*p == ' ' || *p == '\n' || *p == '\r' || *p == '\t')) p++;
  }
  bool lit(const char* w) {
    const size_t n = std::strlen(w);
    if ((size_t)(e - p) >= n && std::strncmp(p, w, n) == 0) {
      p += n;
      return true;
    }
    return false;
  }
  std::string str() {
    std::string o;
    if (p >= e || *p != '"') {
      ok = false;
      return o;
    }
    p++;
    while (p < e && *p != '"') {
      if (*p == '\\' && p + 1 < e) {
        p++;
        switch (*p) {
          case 'n': o += '\n'; break;
          case 't': o += '\t'; break;
          case 'r': o += '\r'; break;
          case 'b': o += '\b'; break;
          case 'f': o += '\f'; break;
          case 'u':
            if (e - p >= 5) {
              o += (char)std::strtol(std::string(p + 1, p + 5).c_str(), nullptr, 16);
              p += 4;
            }
            break;
          default: o += *p;
        }
        p++;
      } else {
        o += *p++;
      }
    }
    if (p >= e) ok = false;
    p++;
    return o;
  }
  JVal val() {
    ws();
    JVal v;
    if (p >= e) {
      ok = false;
      return v;
    }
    if (*p == '{') {
      p++;
      v.k = JVal::OBJ;
      ws();
      if (p < e && *p == '}') {
        p++;
        return v;
      }
      while (ok) {
        ws();
        std::string k = str();
        ws();
        if (p >= e || *p != ':') {
          ok = false;
          break;
        }
        p++;
        v.obj.emplace_back(k, val());
        ws();
        if (p < e && *p == ',') {
          p++;
          continue;
        }
        if (p < e && *p == '}') {
          p++;
          break;
        }
        ok = false;
      }
      return v;
    }
    if (*p == '[') {
      p++;
      v.k = JVal::ARR;
      ws();
      if (p < e && *p == ']') {
        p++;
        return v;
      }
      while (ok) {
        v.arr.push_back(val());
        ws();
        if (p < e && *p == ',') {
          p++;
          continue;
        }
        if (p < e && *p == ']') {
          p++;
          break;
        }
        ok = false;
      }
      return v;
    }
    if (*p == '"') {
      v.k = JVal::STR;
      v.s = str();
      return v;
    }
    if (lit("true")) {
      v.k = JVal::BOOL;
      v.b = true;
      return v;
    }
    if (lit("false")) {
      v.k = JVal::BOOL;
      return v;
    }
    if (lit("null")) return v;
    const char* q = p;
    bool isreal = false;
    while (q < e && (std::strchr("+-0123456789.eEdD", *q))) {
      if (std::strchr(".eEdD", *q)) isreal = true;
      q++;
    }
    std::string num(p, q);
    p = q;
    if (num.empty()) {
      ok = false;
      return v;
    }
    if (isreal) {
      for (char& ch : num)
        if (ch == 'd' || ch == 'D') ch = 'E';
      v.k = JVal::REAL;
      v.d = std::strtod(num.c_str(), nullptr);
    } else {
      v.k = JVal::INT;
      v.i = std::strtoll(num.c_str(), nullptr, 10);
    }
    return v;
  }
};

inline bool read_file(const std::string& fn, std::string& out) {
  FILE* f = std::fopen(fn.c_str(), "rb");
  if (!f) return false;
  char buf[65536];
  size_t n;
  while ((n = std::fread(buf, 1, sizeof buf, f)) > 0) out.append(buf, n);
  std::fclose(f);
  return true;
}

}  // namespace json
}  // namespace wpic
