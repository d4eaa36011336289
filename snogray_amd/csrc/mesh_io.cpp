// Mesh files: PLY (ASCII and binary), Wavefront OBJ, STL (ASCII), MSH and
// UG, read into the vertex / triangle / vertex-normal arrays of an sgr_mesh.
//
//   PLY (load/mesh/load-ply.cc): element "vertex" with x, y, z and
//       optional nx, ny, nz; element "face" with a list property
//       "vertex_indices" of triangles (a longer face is an error, as in the
//       reference); other elements and properties are skipped.
//   OBJ (load/mesh/obj.lua): "v x y z", "vn x y z", "f" with 1-based
//       vertex indices given as "i" or "i//n" (n must equal i), polygons
//       fanned into triangles (v1, v[k-1], v[k]); mtllib, usemtl, g, o and
//       comment lines are ignored.
//   STL (load/mesh/stl.lua): the ASCII grammar of that loader, solids
//       headed "solid ascii"; vertices with equal positions merged (a
//       vertex group, surface/mesh.cc:46-58); facets fanned from their
//       first vertex.
//   MSH (load/mesh/load-msh.cc): sections "<nv> <nt> vertices ...
//       triangles ... [texcoords ...] [normals ...]", triangle indices
//       relative to the section's first vertex, normals made unit length.
//   UG (load/mesh/ug.lua): "v name x y z;", "c name r g b;" (a lambert
//       material) and "f (v1 v2 v3 ...) [material];" lines, {comments};
//       one part per material in order of first use (ug.lua:75-109).
// PLY, OBJ and MSH mark the mesh right-handed (mesh.left_handed = false);
// the STL and UG loaders leave the Mesh default, left-handed
// (surface/mesh.h:55).
#include <algorithm>
#include <array>
#include <cctype>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iterator>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "sgr_internal.h"

struct sgr_mesh_file {
  std::vector<float> v, n;
  std::vector<uint32_t> t;
  std::vector<uint32_t> tri_part;  // empty: every triangle in part 0
  std::vector<float> part_color;   // 3 per part; -1: the caller's material
  int left_handed = 0;
};

namespace {

using sgr::fail;

struct ParseError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

// ---- PLY --------------------------------------------------------------
enum PType { P_I8, P_U8, P_I16, P_U16, P_I32, P_U32, P_F32, P_F64, P_BAD };

PType ptype(const std::string &s) {
  if (s == "char" || s == "int8") return P_I8;
  if (s == "uchar" || s == "uint8") return P_U8;
  if (s == "short" || s == "int16") return P_I16;
  if (s == "ushort" || s == "uint16") return P_U16;
  if (s == "int" || s == "int32") return P_I32;
  if (s == "uint" || s == "uint32") return P_U32;
  if (s == "float" || s == "float32") return P_F32;
  if (s == "double" || s == "float64") return P_F64;
  return P_BAD;
}
size_t psize(PType t) {
  static const size_t sz[] = {1, 1, 2, 2, 4, 4, 4, 8};
  return sz[t];
}

struct Prop {
  std::string name;
  PType type = P_BAD, count_type = P_BAD;
  bool list = false;
};
struct Element {
  std::string name;
  uint64_t count = 0;
  std::vector<Prop> props;
};

struct PlyIn {
  std::istream &in;
  int format;  // 0 ascii, 1 little-endian, 2 big-endian
  bool host_le;
  PlyIn(std::istream &i, int f) : in(i), format(f) {
    const uint32_t x = 1;
    unsigned char c;
    std::memcpy(&c, &x, 1);
    host_le = c == 1;
  }
  double value(PType t) {
    if (format == 0) {
      double d;
      if (!(in >> d)) throw ParseError("PLY: truncated or malformed ASCII data");
      return d;
    }
    unsigned char b[8];
    const size_t n = psize(t);
    if (!in.read((char *)b, (std::streamsize)n)) throw ParseError("PLY: truncated binary data");
    if ((format == 1) != host_le)
      for (size_t i = 0; i < n / 2; i++) std::swap(b[i], b[n - 1 - i]);
    switch (t) {
      case P_I8: { int8_t v; std::memcpy(&v, b, 1); return v; }
      case P_U8: return b[0];
      case P_I16: { int16_t v; std::memcpy(&v, b, 2); return v; }
      case P_U16: { uint16_t v; std::memcpy(&v, b, 2); return v; }
      case P_I32: { int32_t v; std::memcpy(&v, b, 4); return v; }
      case P_U32: { uint32_t v; std::memcpy(&v, b, 4); return v; }
      case P_F32: { float v; std::memcpy(&v, b, 4); return v; }
      case P_F64: { double v; std::memcpy(&v, b, 8); return v; }
      default: throw ParseError("PLY: bad property type");
    }
  }
};

void read_ply(std::istream &in, sgr_mesh_file &m) {
  std::string line;
  if (!std::getline(in, line) || line.substr(0, 3) != "ply")
    throw ParseError("not a PLY file");
  int format = -1;
  std::vector<Element> elems;
  for (;;) {
    if (!std::getline(in, line)) throw ParseError("PLY: header without end_header");
    if (!line.empty() && line.back() == '\r') line.pop_back();
    std::istringstream ls(line);
    std::string kw;
    ls >> kw;
    if (kw == "end_header") break;
    if (kw == "format") {
      std::string f;
      ls >> f;
      format = f == "ascii" ? 0 : f == "binary_little_endian" ? 1 : f == "binary_big_endian" ? 2 : -1;
      if (format < 0) throw ParseError("PLY: unknown format " + f);
    } else if (kw == "element") {
      Element e;
      ls >> e.name >> e.count;
      if (!ls) throw ParseError("PLY: malformed element line");
      elems.push_back(e);
    } else if (kw == "property") {
      if (elems.empty()) throw ParseError("PLY: property before any element");
      Prop p;
      std::string t;
      ls >> t;
      if (t == "list") {
        std::string ct, it;
        ls >> ct >> it >> p.name;
        p.list = true;
        p.count_type = ptype(ct);
        p.type = ptype(it);
        if (p.count_type == P_BAD) throw ParseError("PLY: bad list count type " + ct);
      } else {
        p.type = ptype(t);
        ls >> p.name;
      }
      if (p.type == P_BAD || !ls) throw ParseError("PLY: bad property line: " + line);
      elems.back().props.push_back(p);
    } else if (kw == "comment" || kw == "obj_info" || kw.empty()) {
    } else {
      throw ParseError("PLY: unknown header line: " + line);
    }
  }
  if (format < 0) throw ParseError("PLY: no format line");
  PlyIn rd(in, format);
  bool normals = false;
  for (const Element &e : elems) {
    if (e.name == "vertex") {
      int ix[6] = {-1, -1, -1, -1, -1, -1};
      static const char *names[6] = {"x", "y", "z", "nx", "ny", "nz"};
      for (size_t k = 0; k < e.props.size(); k++)
        for (int c = 0; c < 6; c++)
          if (e.props[k].name == names[c] && !e.props[k].list) ix[c] = (int)k;
      if (ix[0] < 0 || ix[1] < 0 || ix[2] < 0) throw ParseError("PLY: vertex without x, y, z");
      normals = ix[3] >= 0 && ix[4] >= 0 && ix[5] >= 0;
      m.v.reserve(3 * e.count);
      if (normals) m.n.reserve(3 * e.count);
      std::vector<double> val(e.props.size());
      for (uint64_t i = 0; i < e.count; i++) {
        for (size_t k = 0; k < e.props.size(); k++) {
          const Prop &p = e.props[k];
          if (p.list) {
            const uint64_t c = (uint64_t)rd.value(p.count_type);
            for (uint64_t q = 0; q < c; q++) rd.value(p.type);
          } else {
            val[k] = rd.value(p.type);
          }
        }
        for (int c = 0; c < 3; c++) m.v.push_back((float)val[ix[c]]);
        if (normals)
          for (int c = 3; c < 6; c++) m.n.push_back((float)val[ix[c]]);
      }
    } else if (e.name == "face") {
      int fi = -1;
      for (size_t k = 0; k < e.props.size(); k++)
        if (e.props[k].list &&
            (e.props[k].name == "vertex_indices" || e.props[k].name == "vertex_index"))
          fi = (int)k;
      if (fi < 0) throw ParseError("PLY: face without vertex_indices");
      m.t.reserve(3 * e.count);
      for (uint64_t i = 0; i < e.count; i++)
        for (size_t k = 0; k < e.props.size(); k++) {
          const Prop &p = e.props[k];
          if (p.list) {
            const uint64_t c = (uint64_t)rd.value(p.count_type);
            if ((int)k == fi && c != 3) throw ParseError("Invalid number of indices in face");
            for (uint64_t q = 0; q < c; q++) {
              const double x = rd.value(p.type);
              if ((int)k == fi) {
                if (x < 0) throw ParseError("PLY: negative vertex index");
                m.t.push_back((uint32_t)x);
              }
            }
          } else {
            rd.value(p.type);
          }
        }
    } else {  // skipped element
      for (uint64_t i = 0; i < e.count; i++)
        for (const Prop &p : e.props) {
          if (p.list) {
            const uint64_t c = (uint64_t)rd.value(p.count_type);
            for (uint64_t q = 0; q < c; q++) rd.value(p.type);
          } else {
            rd.value(p.type);
          }
        }
    }
  }
}

// ---- OBJ --------------------------------------------------------------
void read_obj(std::istream &in, sgr_mesh_file &m) {
  std::string line;
  size_t lineno = 0;
  auto err = [&](const std::string &msg) {
    throw ParseError("OBJ line " + std::to_string(lineno) + ": " + msg);
  };
  while (std::getline(in, line)) {
    lineno++;
    if (!line.empty() && line.back() == '\r') line.pop_back();
    std::istringstream ls(line);
    std::string cmd;
    if (!(ls >> cmd) || cmd[0] == '#' || cmd == "g" || cmd == "o" || cmd == "usemtl" ||
        cmd == "mtllib" || cmd == "s")
      continue;
    if (cmd == "v" || cmd == "vn") {
      float x, y, z;
      if (!(ls >> x >> y >> z)) err("expected three numbers");
      auto &dst = cmd == "v" ? m.v : m.n;
      dst.push_back(x);
      dst.push_back(y);
      dst.push_back(z);
    } else if (cmd == "f") {
      std::vector<uint32_t> idx;
      std::string tok;
      while (ls >> tok) {
        const size_t sl = tok.find("//");
        long vi, ni = -1;
        char *end;
        vi = std::strtol(tok.c_str(), &end, 10);
        if (sl == std::string::npos) {
          if (*end) err("vertex index '" + tok + "' (only i and i//n are read)");
        } else {
          if (end != tok.c_str() + sl) err("vertex index '" + tok + "'");
          ni = std::strtol(tok.c_str() + sl + 2, &end, 10);
          if (*end) err("vertex index '" + tok + "'");
          if (ni != vi) err("Normal indices must be identical to vertex indices");
        }
        if (vi < 1) err("vertex index '" + tok + "' (indices start at 1)");
        idx.push_back((uint32_t)(vi - 1));
      }
      if (idx.size() < 3) err("a face needs three vertices");
      for (size_t k = 2; k < idx.size(); k++) {  // fan: (v1, prev, next)
        m.t.push_back(idx[0]);
        m.t.push_back(idx[k - 1]);
        m.t.push_back(idx[k]);
      }
    } else {
      err("unknown command '" + cmd + "'");
    }
  }
  if (!m.n.empty() && m.n.size() != m.v.size())
    throw ParseError("OBJ: " + std::to_string(m.n.size() / 3) + " normals for " +
                     std::to_string(m.v.size() / 3) + " vertices");
}

// ---- text scanning for the PEG grammars of the Lua loaders -----------
// lua-util/lpeg-utils.lua: FLOAT = [+-]? ("." digits | digits ("." digits?)?)
// ("e" [+-]? digits)?, converted by tonumber (a double) and stored as a
// float coordinate.
struct Scan {
  const std::string &s;
  explicit Scan(const std::string &t) : s(t) {}
  bool at(size_t p, char c) const { return p < s.size() && s[p] == c; }
  bool digit(size_t p) const { return p < s.size() && s[p] >= '0' && s[p] <= '9'; }
  bool lit(size_t &p, const char *w) const {
    const size_t n = std::strlen(w);
    if (s.compare(p, n, w) != 0) return false;
    p += n;
    return true;
  }
  void hws(size_t &p) const {  // OPT_HORIZ_WS
    while (at(p, ' ') || at(p, '\t')) p++;
  }
  void ws(size_t &p) const {  // OPT_WS
    while (p < s.size() && (s[p] == ' ' || s[p] == '\t' || s[p] == '\r' || s[p] == '\n' ||
                            s[p] == '\f'))
      p++;
  }
  bool flt(size_t &p, double &out) const {
    size_t q = p;
    if (at(q, '+') || at(q, '-')) q++;
    if (at(q, '.') && digit(q + 1)) {
      q++;
      while (digit(q)) q++;
    } else if (digit(q)) {
      while (digit(q)) q++;
      if (at(q, '.')) {
        q++;
        while (digit(q)) q++;
      }
    } else {
      return false;
    }
    size_t e = q;
    if (at(e, 'e')) {
      e++;
      if (at(e, '+') || at(e, '-')) e++;
      if (digit(e)) {
        while (digit(e)) e++;
        q = e;
      }
    }
    out = std::strtod(s.substr(p, q - p).c_str(), nullptr);
    p = q;
    return true;
  }
  size_t line(size_t p) const {
    return 1 + (size_t)std::count(s.begin(), s.begin() + (ptrdiff_t)std::min(p, s.size()), '\n');
  }
};

std::string slurp(std::istream &in) {
  return std::string(std::istreambuf_iterator<char>(in), std::istreambuf_iterator<char>());
}

// ---- STL (ASCII) -------------------------------------------------------
// The grammar of load/mesh/stl.lua:
//   solid  = "solid" hws "ascii" (ws facet)* ws "end" hws "solid" ws
//   facet  = "facet" (ws (loop | "normal" coords))* ws "end" hws "facet"
//   loop   = "outer" hws "loop" (ws "vertex" coords)* ws "end" hws "loop"
//   coords = (hws float){3}
// matched repeatedly until the text ends.  Every partial match of a facet
// leads the whole match to fail, so the parser stops at the first
// mismatch.  Vertices go through one vertex group for the file (equal
// positions share an index, surface/mesh.cc:46-58); a facet of n vertices
// adds the fan (v1, v[k-1], v[k]).
void read_stl(const std::string &text, sgr_mesh_file &m) {
  const Scan S(text);
  struct Less {
    bool operator()(const std::array<float, 3> &a, const std::array<float, 3> &b) const {
      return a[0] < b[0] || (a[0] == b[0] && (a[1] < b[1] || (a[1] == b[1] && a[2] < b[2])));
    }
  };
  std::map<std::array<float, 3>, uint32_t, Less> group;
  std::vector<uint32_t> fv;
  size_t sync = 0;
  auto err = [&](const std::string &msg) {
    throw ParseError("STL line " + std::to_string(S.line(sync)) + ": " + msg);
  };
  auto coords = [&](size_t &p, double *x) {
    for (int k = 0; k < 3; k++) {
      S.hws(p);
      if (!S.flt(p, x[k])) err("parse error");
    }
  };
  auto end_kw = [&](size_t &p, const char *kw) {
    S.ws(p);
    sync = p;
    if (!S.lit(p, "end")) err("parse error");
    S.hws(p);
    if (!S.lit(p, kw)) err("parse error");
  };
  size_t pos = 0;
  while (pos < text.size()) {
    size_t p = pos;
    sync = p;
    if (!S.lit(p, "solid")) err("parse error");
    S.hws(p);
    if (!S.lit(p, "ascii")) err("parse error");
    for (;;) {
      size_t q = p;
      S.ws(q);
      sync = q;
      if (!S.lit(q, "facet")) break;
      for (;;) {
        size_t r = q;
        S.ws(r);
        sync = r;
        if (S.lit(r, "outer")) {
          S.hws(r);
          if (!S.lit(r, "loop")) err("parse error");
          for (;;) {
            size_t u = r;
            S.ws(u);
            sync = u;
            if (!S.lit(u, "vertex")) break;
            double x[3];
            coords(u, x);
            const std::array<float, 3> key{(float)x[0], (float)x[1], (float)x[2]};
            auto it = group.find(key);
            if (it == group.end()) {
              it = group.emplace(key, (uint32_t)(m.v.size() / 3)).first;
              m.v.insert(m.v.end(), key.begin(), key.end());
            }
            fv.push_back(it->second);
            r = u;
          }
          end_kw(r, "loop");
        } else if (S.lit(r, "normal")) {
          double x[3];
          coords(r, x);
        } else {
          break;
        }
        q = r;
      }
      end_kw(q, "facet");
      if (fv.size() < 3) err("not enough vertices in facet");
      for (size_t k = 2; k < fv.size(); k++) {
        m.t.push_back(fv[0]);
        m.t.push_back(fv[k - 1]);
        m.t.push_back(fv[k]);
      }
      fv.clear();
      p = q;
    }
    end_kw(p, "solid");
    S.ws(p);
    pos = p;
  }
  m.left_handed = 1;
}

// ---- MSH ----------------------------------------------------------------
// load/mesh/load-msh.cc: whitespace-separated sections
//   <num_vertices> <num_triangles> vertices x y z ... triangles i j k ...
//   [texcoords u v ...] [normals x y z ...]
// until the end of the file; triangle indices count from the section's
// first vertex; normals are made unit length (v * (1 / |v|)).  Numbers are
// read by the stream extractors, as there.  A file whose sections do not
// all carry normals is refused (the reference then computes the missing
// ones by smoothing, surface/mesh.cc:121-137).
void read_msh(std::istream &in, sgr_mesh_file &m) {
  auto bad = [](const std::string &msg) { throw ParseError("MSH: " + msg); };
  std::string kw;
  if (!(in >> kw)) bad("empty file");
  size_t sections = 0, with_normals = 0;
  do {
    const uint32_t base = (uint32_t)(m.v.size() / 3);
    const unsigned nv = (unsigned)std::atoi(kw.c_str());
    unsigned nt;
    if (!(in >> nt) || !(in >> kw) || kw != "vertices") bad("bad format (expected vertices)");
    for (unsigned i = 0; i < 3 * nv; i++) {
      float x;
      if (!(in >> x)) bad("bad format (vertex coordinates)");
      m.v.push_back(x);
    }
    if (!(in >> kw) || kw != "triangles") bad("bad format (expected triangles)");
    for (unsigned i = 0; i < 3 * nt; i++) {
      unsigned ix;
      if (!(in >> ix)) bad("bad format (triangle indices)");
      m.t.push_back(base + ix);
    }
    sections++;
    const bool more = (bool)(in >> kw);
    bool next = more;
    if (next && kw == "texcoords") {
      for (unsigned i = 0; i < 2 * nv; i++) {
        float u;
        if (!(in >> u)) bad("bad format (texcoords)");
      }
      next = (bool)(in >> kw);
    }
    if (next && kw == "normals") {
      for (unsigned i = 0; i < nv; i++) {
        float x, y, z;
        if (!(in >> x >> y >> z)) bad("bad format (normals)");
        const float len = std::sqrt(x * x + y * y + z * z);
        const float inv = 1 / len;
        m.n.push_back(x * inv);
        m.n.push_back(y * inv);
        m.n.push_back(z * inv);
      }
      with_normals++;
      next = (bool)(in >> kw);
    }
    if (!next) {
      if (!in.eof()) bad("bad format");
      break;
    }
  } while (true);
  if (with_normals && with_normals != sections)
    bad("normals in some sections only (not supported)");
  m.left_handed = 0;
}

// ---- UG -----------------------------------------------------------------
// The line grammar of load/mesh/ug.lua, matched once per line:
//   line  = (cmd | hws) nl       nl = "\r"? "\n"
//   cmd   = ("v" name float float float | "c" name float float float
//            | "f" hws "(" name name name name* hws ")" name*) ";" hws
// where every name and float may be preceded by hws, and hws includes
// {comments}.  "v" defines a named vertex (each definition a new vertex),
// "c" a named lambert material, "f" a polygon fanned from its first
// vertex.  Of the names after the third, the last is the face's material
// and the others are fan vertices (so a polygon of four or more vertices
// needs a material name); a face without one takes the caller's material.
// Each material (each "c" definition) gets a part on first use.
void read_ug(const std::string &text, sgr_mesh_file &m) {
  const Scan S(text);
  size_t line_start = 0;
  auto err = [&](const std::string &msg) {
    throw ParseError("UG line " + std::to_string(S.line(line_start)) + ": " + msg);
  };
  auto hws = [&](size_t &p) {  // horizontal whitespace and {comments}
    for (;;) {
      S.hws(p);
      if (!S.at(p, '{')) return;
      const size_t close = text.find('}', p + 1);
      if (close == std::string::npos) return;
      p = close + 1;
    }
  };
  auto alpha = [&](size_t p) {
    const char c = p < text.size() ? text[p] : 0;
    return (c >= 'A' && c <= 'Z') || (c >= 'a' && c <= 'z') || c == '_';
  };
  auto name = [&](size_t &p, std::string &out) {
    size_t q = p;
    hws(q);
    if (!alpha(q)) return false;
    const size_t b = q;
    while (alpha(q) || S.digit(q)) q++;
    out = text.substr(b, q - b);
    p = q;
    return true;
  };
  auto flt = [&](size_t &p, double &x) {
    size_t q = p;
    hws(q);
    if (!S.flt(q, x)) return false;
    p = q;
    return true;
  };
  std::map<std::string, uint32_t> verts, mats;  // name -> vertex / material definition
  std::vector<std::array<float, 3>> mat_color;  // per definition (1-based ids)
  std::map<uint32_t, uint32_t> part_of;         // definition (0: default) -> part
  std::vector<uint32_t> tri_part;
  size_t pos = 0;
  while (pos < text.size()) {
    line_start = pos;
    size_t p = pos;
    // try a command; on any mismatch the line must be blank
    bool cmd = false;
    std::string nm;
    double x[3];
    std::vector<std::string> names;
    char kind = 0;
    {
      size_t q = p;
      if (S.at(q, 'v') || S.at(q, 'c')) {
        kind = text[q++];
        if (name(q, nm) && flt(q, x[0]) && flt(q, x[1]) && flt(q, x[2]) && S.at(q, ';')) {
          q++;
          hws(q);
          cmd = true;
          p = q;
        }
      } else if (S.at(q, 'f')) {
        kind = 'f';
        q++;
        hws(q);
        if (S.at(q, '(')) {
          q++;
          std::string n1;
          while (name(q, n1)) names.push_back(n1);
          size_t r = q;
          hws(r);
          if (names.size() >= 3 && S.at(r, ')')) {
            r++;
            while (name(r, n1)) names.push_back(n1);
            if (S.at(r, ';')) {
              r++;
              hws(r);
              cmd = true;
              p = r;
            }
          }
        }
      }
    }
    if (!cmd) hws(p);
    if (S.at(p, '\r')) p++;
    if (!S.at(p, '\n')) err("parse error");
    pos = p + 1;
    if (!cmd) continue;
    if (kind == 'v') {
      verts[nm] = (uint32_t)(m.v.size() / 3);
      for (int k = 0; k < 3; k++) m.v.push_back((float)x[k]);
    } else if (kind == 'c') {
      mat_color.push_back({(float)x[0], (float)x[1], (float)x[2]});
      mats[nm] = (uint32_t)mat_color.size();
    } else {
      uint32_t def = 0;
      if (names.size() > 3) {
        auto it = mats.find(names.back());
        if (it == mats.end()) err("Unknown material name \"" + names.back() + "\"");
        def = it->second;
      }
      auto pt = part_of.find(def);
      if (pt == part_of.end()) {
        pt = part_of.emplace(def, (uint32_t)(m.part_color.size() / 3)).first;
        for (int k = 0; k < 3; k++) m.part_color.push_back(def ? mat_color[def - 1][k] : -1.0f);
      }
      auto vert = [&](const std::string &vn) {
        auto it = verts.find(vn);
        if (it == verts.end()) err("Unknown vertex name \"" + vn + "\"");
        return it->second;
      };
      const uint32_t v1 = vert(names[0]), v2 = vert(names[1]);
      uint32_t prev = vert(names[2]);
      m.t.push_back(v1);
      m.t.push_back(v2);
      m.t.push_back(prev);
      tri_part.push_back(pt->second);
      for (size_t k = 3; k + 1 < names.size(); k++) {
        const uint32_t vn = vert(names[k]);
        m.t.push_back(v1);
        m.t.push_back(prev);
        m.t.push_back(vn);
        tri_part.push_back(pt->second);
        prev = vn;
      }
    }
  }
  m.tri_part = std::move(tri_part);
  m.left_handed = 1;
}

}  // namespace

extern "C" {

int sgr_mesh_file_open(const char *path, const char *format, sgr_mesh_file **out,
                       uint64_t *num_vertices, uint64_t *num_triangles, int *has_normals) {
  if (!path || !out) return fail(SGR_ERR_INVALID, "sgr_mesh_file_open: null argument");
  *out = nullptr;
  return sgr::guarded("sgr_mesh_file_open", [&] {
    std::string f = format ? format : "";
    if (f.empty()) {
      const char *dot = std::strrchr(path, '.');
      if (dot) f = dot + 1;
    }
    for (auto &c : f) c = (char)std::tolower((unsigned char)c);
    if (f != "ply" && f != "obj" && f != "stl" && f != "msh" && f != "ug")
      return fail(SGR_ERR_INVALID, "mesh file: unknown format '" + f + "'");
    std::ifstream in(path, std::ios::binary);
    if (!in)
      return fail(SGR_ERR_INVALID, std::string("mesh file: cannot open ") + path + ": " +
                                       std::strerror(errno));
    std::unique_ptr<sgr_mesh_file> m(new sgr_mesh_file);
    try {
      if (f == "ply") read_ply(in, *m);
      else if (f == "obj") read_obj(in, *m);
      else if (f == "stl") read_stl(slurp(in), *m);
      else if (f == "msh") read_msh(in, *m);
      else read_ug(slurp(in), *m);
    } catch (const ParseError &e) {
      return fail(SGR_ERR_INVALID, std::string(path) + ": " + e.what());
    }
    const uint64_t nv = m->v.size() / 3;
    if (nv >= (1ull << 32) || m->t.size() / 3 >= (1ull << 32))
      return fail(SGR_ERR_LIMIT, std::string(path) + ": more than 2^32 vertices or triangles");
    for (uint32_t ix : m->t)
      if (ix >= nv)
        return fail(SGR_ERR_INVALID, std::string(path) + ": vertex index " + std::to_string(ix) +
                                         " out of range (" + std::to_string(nv) + " vertices)");
    if (num_vertices) *num_vertices = nv;
    if (num_triangles) *num_triangles = m->t.size() / 3;
    if (has_normals) *has_normals = m->n.empty() ? 0 : 1;
    *out = m.release();
    return SGR_OK;
  });
}

int sgr_mesh_file_copy(const sgr_mesh_file *mf, float *vertices, uint32_t *triangles,
                       float *normals) {
  if (!mf) return fail(SGR_ERR_INVALID, "sgr_mesh_file_copy: null handle");
  if (vertices && !mf->v.empty()) std::memcpy(vertices, mf->v.data(), mf->v.size() * 4);
  if (triangles && !mf->t.empty()) std::memcpy(triangles, mf->t.data(), mf->t.size() * 4);
  if (normals && !mf->n.empty()) std::memcpy(normals, mf->n.data(), mf->n.size() * 4);
  return SGR_OK;
}

int sgr_mesh_file_parts(const sgr_mesh_file *mf, int *left_handed, uint32_t *num_parts,
                        uint32_t *tri_parts, float *part_colors) {
  if (!mf) return fail(SGR_ERR_INVALID, "sgr_mesh_file_parts: null handle");
  const size_t nt = mf->t.size() / 3;
  // formats without materials load into one part, the caller's material
  const bool one = mf->part_color.empty();
  const uint32_t np = one ? (nt ? 1u : 0u) : (uint32_t)(mf->part_color.size() / 3);
  if (left_handed) *left_handed = mf->left_handed;
  if (num_parts) *num_parts = np;
  if (tri_parts && nt) {
    if (one) std::memset(tri_parts, 0, nt * 4);
    else std::memcpy(tri_parts, mf->tri_part.data(), nt * 4);
  }
  if (part_colors && np) {
    if (one) part_colors[0] = part_colors[1] = part_colors[2] = -1.0f;
    else std::memcpy(part_colors, mf->part_color.data(), mf->part_color.size() * 4);
  }
  return SGR_OK;
}

int sgr_mesh_file_close(sgr_mesh_file *mf) {
  delete mf;
  return SGR_OK;
}

}  // extern "C"
