// Mesh files: PLY (ASCII and binary) and Wavefront OBJ, read into the
// vertex / triangle / vertex-normal arrays of an sgr_mesh.
//
//   PLY (load/mesh/load-ply.cc): element "vertex" with x, y, z and
//       optional nx, ny, nz; element "face" with a list property
//       "vertex_indices" of triangles (a longer face is an error, as in the
//       reference); other elements and properties are skipped.
//   OBJ (load/mesh/obj.lua): "v x y z", "vn x y z", "f" with 1-based
//       vertex indices given as "i" or "i//n" (n must equal i), polygons
//       fanned into triangles (v1, v[k-1], v[k]); mtllib, usemtl, g, o and
//       comment lines are ignored.
// Both loaders mark the mesh right-handed (mesh.left_handed = false).
#include <cctype>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "sgr_internal.h"

struct sgr_mesh_file {
  std::vector<float> v, n;
  std::vector<uint32_t> t;
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
    if (f != "ply" && f != "obj")
      return fail(SGR_ERR_INVALID, "mesh file: unknown format '" + f + "'");
    std::ifstream in(path, std::ios::binary);
    if (!in)
      return fail(SGR_ERR_INVALID, std::string("mesh file: cannot open ") + path + ": " +
                                       std::strerror(errno));
    std::unique_ptr<sgr_mesh_file> m(new sgr_mesh_file);
    try {
      if (f == "ply") read_ply(in, *m);
      else read_obj(in, *m);
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

int sgr_mesh_file_close(sgr_mesh_file *mf) {
  delete mf;
  return SGR_OK;
}

}  // extern "C"
