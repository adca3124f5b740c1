// surface.cpp -- host side of the surface contract: OBJ ingestion
// (Surface::init_from_obj, src/Utils/Surface.cpp:57-133) and PLY export
// (Surface::export_to_ply, src/Utils/Surface.cpp:138-285).  Plain C++: file
// input and output around the GPU path, no device code.
//
// All float arithmetic here mirrors the reference's glm::vec3 float
// operations term by term, so no expression may be contracted into fused
// multiply-adds.
#pragma clang fp contract(off)

#include <algorithm>
#include <charconv>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>
#include <unordered_map>
#include <vector>

#include "femhip_step.h"

struct fem_surface {
  std::vector<float> xyz;       // [nv][3]
  std::vector<float> normals;   // [nv][3]
  std::vector<uint32_t> tri;    // [T][3]
  std::vector<uint8_t> onb;     // [nv]
  int32_t num_boundary = 0;
};

namespace femhip {

namespace {

struct Corner {
  int v, n;  // vertex and normal index (0-based; n = -1 without a normal)
};

// tinyobjloader's index rule: 1-based, negative counts back from the end of
// the list read so far, 0 invalid.
bool fix_index(long raw, size_t count, int* out) {
  if (raw > 0) {
    *out = (int)(raw - 1);
    return true;
  }
  if (raw < 0 && (size_t)(-raw) <= count) {
    *out = (int)((long)count + raw);
    return true;
  }
  return false;
}

// "v", "v/t", "v//n" or "v/t/n".
bool parse_corner(const char* tok, size_t nv, size_t nn, Corner* c) {
  char* end = nullptr;
  const long v = std::strtol(tok, &end, 10);
  if (end == tok || !fix_index(v, nv, &c->v)) return false;
  c->n = -1;
  if (*end != '/') return true;
  const char* t = end + 1;
  if (*t != '/') std::strtol(t, &end, 10);  // texture index: not used
  else end = (char*)t;
  if (*end != '/') return true;
  t = end + 1;
  const long n = std::strtol(t, &end, 10);
  if (end == t) return true;
  return fix_index(n, nn, &c->n);
}

// Split of a face into triangles (the reader's triangulate = true): a
// triangle as it is; a quad along its shorter diagonal, lengths in float as
// tinyobjloader computes them ([0,1,2],[0,2,3] when |v2 - v0| < |v3 - v1|,
// else [0,1,3],[1,2,3]); larger polygons as a fan around the first corner.
void triangulate(const std::vector<Corner>& f, const std::vector<float>& V, std::vector<Corner>& out) {
  auto push = [&](int a, int b, int c) {
    out.push_back(f[a]);
    out.push_back(f[b]);
    out.push_back(f[c]);
  };
  if (f.size() == 3) {
    push(0, 1, 2);
  } else if (f.size() == 4) {
    const float* p0 = &V[3 * (size_t)f[0].v];
    const float* p1 = &V[3 * (size_t)f[1].v];
    const float* p2 = &V[3 * (size_t)f[2].v];
    const float* p3 = &V[3 * (size_t)f[3].v];
    const float ax = p2[0] - p0[0], ay = p2[1] - p0[1], az = p2[2] - p0[2];
    const float bx = p3[0] - p1[0], by = p3[1] - p1[1], bz = p3[2] - p1[2];
    const float s02 = ax * ax + ay * ay + az * az;
    const float s13 = bx * bx + by * by + bz * bz;
    if (s02 < s13) {
      push(0, 1, 2);
      push(0, 2, 3);
    } else {
      push(0, 1, 3);
      push(1, 2, 3);
    }
  } else {
    for (size_t q = 1; q + 1 < f.size(); q++) push(0, (int)q, (int)q + 1);
  }
}

int read_obj(const char* path, fem_surface* s) {
  std::ifstream in(path);
  if (!in) {
    set_error("A valid file was not provided.");
    return FEM_ERR_INVALID;
  }
  std::vector<float> V, N;
  std::vector<std::vector<Corner>> shapes(1);  // triangulated corners per shape
  std::vector<Corner> face;
  std::string line;
  long lineno = 0;
  while (std::getline(in, line)) {
    lineno++;
    if (!line.empty() && line.back() == '\r') line.pop_back();
    const char* p = line.c_str();
    while (*p == ' ' || *p == '\t') p++;
    if (*p == '\0' || *p == '#') continue;
    auto tag_is = [&](const char* t) {
      const size_t l = std::strlen(t);
      return std::strncmp(p, t, l) == 0 && (p[l] == ' ' || p[l] == '\t' || p[l] == '\0');
    };
    if (tag_is("v") || tag_is("vn")) {
      const bool is_n = p[1] == 'n';
      const char* q = p + (is_n ? 2 : 1);
      for (int k = 0; k < 3; k++) {
        char* end = nullptr;
        const double d = std::strtod(q, &end);
        if (end == q) {
          set_error(std::string(path) + ":" + std::to_string(lineno) + ": malformed " + (is_n ? "vn" : "v") + " line");
          return FEM_ERR_INVALID;
        }
        (is_n ? N : V).push_back((float)d);
        q = end;
      }
    } else if (tag_is("f")) {
      face.clear();
      const char* q = p + 1;
      while (true) {
        while (*q == ' ' || *q == '\t') q++;
        if (*q == '\0') break;
        const char* tok = q;
        while (*q && *q != ' ' && *q != '\t') q++;
        const std::string t(tok, q - tok);
        Corner c;
        if (!parse_corner(t.c_str(), V.size() / 3, N.size() / 3, &c) || (size_t)c.v >= V.size() / 3) {
          set_error(std::string(path) + ":" + std::to_string(lineno) + ": face index out of range");
          return FEM_ERR_INVALID;
        }
        face.push_back(c);
      }
      if (face.size() >= 3) triangulate(face, V, shapes.back());
    } else if (tag_is("o") || tag_is("g")) {
      if (!shapes.back().empty()) shapes.emplace_back();
    }
  }
  if (shapes.size() > 1 && shapes.back().empty()) shapes.pop_back();
  if (V.empty()) {
    set_error(std::string("File ") + path + " does not contain a valid .obj mesh.");
    return FEM_ERR_INVALID;
  }
  if (N.empty()) {
    set_error(std::string("File ") + path + " does not contain normals!");
    return FEM_ERR_INVALID;
  }
  const size_t nshapes = shapes.size();
  // Every shape appends the whole vertex list again; triangle indices keep
  // pointing at the first copy (Surface.cpp:77-80).
  s->xyz.clear();
  for (size_t q = 0; q < nshapes; q++) s->xyz.insert(s->xyz.end(), V.begin(), V.end());
  s->tri.clear();
  for (const auto& sh : shapes)
    for (const Corner& c : sh) s->tri.push_back((uint32_t)c.v);
  const size_t nv = s->xyz.size() / 3;
  // Vertex normals: corner normals summed per vertex in file order, then
  // glm::normalize (v * 1/sqrt(dot(v, v)); a vertex without corners gets
  // 0 * inf = NaN, as the reference's) (Surface.cpp:89-102).
  s->normals.assign(3 * nv, 0.0f);
  for (const auto& sh : shapes)
    for (const Corner& c : sh) {
      if (c.n < 0 || (size_t)c.n >= N.size() / 3) continue;  // corner without a normal
      for (int k = 0; k < 3; k++) s->normals[3 * (size_t)c.v + k] += N[3 * (size_t)c.n + k];
    }
  for (size_t i = 0; i < nv; i++) {
    float* n = &s->normals[3 * i];
    const float d = n[0] * n[0] + n[1] * n[1] + n[2] * n[2];
    const float inv = 1.0f / std::sqrt(d);
    for (int k = 0; k < 3; k++) n[k] = n[k] * inv;
  }
  // Boundary: endpoints of every edge used by exactly one triangle
  // (Surface.cpp:104-127).
  std::unordered_map<uint64_t, int> edges;
  edges.reserve(s->tri.size() * 2);
  for (size_t t = 0; t + 2 < s->tri.size(); t += 3)
    for (int j = 0; j < 3; j++) {
      const uint64_t a = s->tri[t + j], b = s->tri[t + (j + 1) % 3];
      edges[(std::min(a, b) << 32) | std::max(a, b)]++;
    }
  s->onb.assign(nv, 0);
  for (const auto& e : edges)
    if (e.second == 1) {
      s->onb[e.first >> 32] = 1;
      s->onb[e.first & 0xffffffffu] = 1;
    }
  s->num_boundary = 0;
  for (uint8_t b : s->onb) s->num_boundary += b;
  return 0;
}

// ---- PLY export ------------------------------------------------------------

struct V3 {
  float x, y, z;
};
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator*(float s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
// glm::mix(x, y, a) = x * (1 - a) + y * a
inline V3 mix(V3 x, V3 y, float a) {
  const float b = 1.0f - a;
  return {x.x * b + y.x * a, x.y * b + y.y * a, x.z * b + y.z * a};
}

// Position identity of the export's vertex map: glm::vec3 equality (so -0.0
// equals 0.0) and a hash that agrees with it.
struct V3Eq {
  bool operator()(const V3& a, const V3& b) const { return a.x == b.x && a.y == b.y && a.z == b.z; }
};
struct V3Hash {
  size_t operator()(const V3& v) const {
    auto h = [](float f) {
      uint32_t u;
      f = (f == 0.0f) ? 0.0f : f;
      std::memcpy(&u, &f, 4);
      return (size_t)u;
    };
    size_t s = h(v.x);
    s ^= h(v.y) + 0x9e3779b97f4a7c15ull + (s << 6) + (s >> 2);
    s ^= h(v.z) + 0x9e3779b97f4a7c15ull + (s << 6) + (s >> 2);
    return s;
  }
};

// The reference's default colour map, Viridis as an order-6 polynomial per
// channel (src/Utils/AssetStorage.cpp:141-152), evaluated as
// ColorMap::get_color (src/Utils/ColorMap.cpp:12-14).
const float kViridis[7][3] = {
    {0.274344f, 0.004462f, 0.331359f},   {0.108915f, 1.397291f, 1.388110f},  {-0.319631f, 0.243490f, 0.156419f},
    {-4.629188f, -5.882803f, -19.646115f}, {6.181719f, 14.388598f, 57.442181f}, {4.876952f, -13.955112f, -66.125783f},
    {-5.513165f, 4.709245f, 26.582180f}};

V3 viridis(float t) {
  float c[3];
  for (int k = 0; k < 3; k++) {
    float acc = kViridis[6][k];
    for (int i = 5; i >= 0; i--) acc = kViridis[i][k] + t * acc;
    c[k] = acc;
  }
  return {c[0], c[1], c[2]};
}

// std::format("{}") of a float: the shortest text that reads back exactly.
void put_float(std::string& out, float f) {
  char buf[64];
  const auto r = std::to_chars(buf, buf + sizeof(buf), f);
  out.append(buf, r.ptr);
}

}  // namespace

int ply_export(const float* xyz, int32_t nv, const uint32_t* tri, int32_t ntri, const float* normals,
               const float* values, float vertex_extrusion, float threshold, int32_t mesh_type, const char* path) {
  if (mesh_type < FEM_MESH_OPEN || mesh_type > FEM_MESH_MIRRORED) {
    set_error("export_ply: unknown mesh type");
    return FEM_ERR_INVALID;
  }
  for (int32_t t = 0; t < 3 * ntri; t++)
    if ((int64_t)tri[t] >= nv) {
      set_error("export_ply: triangle index out of range");
      return FEM_ERR_INVALID;
    }
  auto P = [&](uint32_t i) { return V3{xyz[3 * (size_t)i], xyz[3 * (size_t)i + 1], xyz[3 * (size_t)i + 2]}; };
  auto Nrm = [&](uint32_t i) { return V3{normals[3 * (size_t)i], normals[3 * (size_t)i + 1], normals[3 * (size_t)i + 2]}; };
  std::vector<V3> pos;
  std::vector<float> pval;
  std::vector<uint32_t> faces;
  std::unordered_map<V3, int, V3Hash, V3Eq> map;
  // add_clipped_position, then the index the reference reads back with
  // operator[]: a position with a NaN component never compares equal, so it
  // is appended and then read back as index 0
  auto add = [&](V3 p, float value) {
    if (map.find(p) == map.end()) {
      map.emplace(p, (int)pos.size());
      pos.push_back(p);
      pval.push_back(value);
    }
  };
  auto index_of = [&](V3 p) -> uint32_t {
    auto it = map.find(p);
    return it == map.end() ? 0u : (uint32_t)it->second;
  };
  const float ext = std::max(0.0f, vertex_extrusion);
  auto extruded = [&](uint32_t i) { return (values[i] * ext) * Nrm(i) + P(i); };
  auto projected = [&](uint32_t i) {
    const float sc = mesh_type == FEM_MESH_MIRRORED ? (threshold - values[i] * vertex_extrusion)
                                                    : (threshold * vertex_extrusion);
    return sc * Nrm(i) + P(i);
  };
  auto face = [&](uint32_t a, uint32_t b, uint32_t c) {
    faces.push_back(a);
    faces.push_back(b);
    faces.push_back(c);
  };
  for (int32_t t = 0; t < ntri; t++) {
    const uint32_t T[3] = {tri[3 * (size_t)t], tri[3 * (size_t)t + 1], tri[3 * (size_t)t + 2]};
    bool above[3];
    for (int i = 0; i < 3; i++) above[i] = (double)(values[T[i]] - threshold) >= 1e-9;
    const int count = above[0] + above[1] + above[2];
    if (count == 1) {
      const int ai = above[0] ? 0 : above[1] ? 1 : 2;
      uint32_t clipped[3];
      for (int i = 0; i < 3; i++) {
        V3 p;
        float v;
        if (i != ai) {
          p = mix(extruded(T[i]), extruded(T[ai]), (threshold - values[T[i]]) / (values[T[ai]] - values[T[i]]));
          v = threshold;
        } else {
          p = extruded(T[i]);
          v = values[T[ai]];
        }
        add(p, v);
        clipped[i] = index_of(p);
      }
      face(clipped[0], clipped[1], clipped[2]);
      if (mesh_type != FEM_MESH_OPEN) {
        uint32_t pr[3];
        for (int i = 0; i < 3; i++) {
          if (i != ai) {
            pr[i] = clipped[i];
          } else {
            const V3 np = projected(T[i]);
            add(np, threshold);
            pr[i] = index_of(np);
          }
        }
        face(pr[0], pr[1], pr[2]);
      }
    } else if (count == 2) {
      const int a1 = above[0] ? 0 : 1;
      const int a2 = above[0] && above[1] ? 1 : 2;
      const int bi = 3 - (a1 + a2);
      const V3 p1 = extruded(T[a1]), p2 = extruded(T[a2]);
      const V3 q1 = mix(extruded(T[bi]), p1, (threshold - values[T[bi]]) / (values[T[a1]] - values[T[bi]]));
      const V3 q2 = mix(extruded(T[bi]), p2, (threshold - values[T[bi]]) / (values[T[a2]] - values[T[bi]]));
      add(p1, values[T[a1]]);
      add(p2, values[T[a2]]);
      add(q1, threshold);
      add(q2, threshold);
      face(index_of(q1), index_of(p1), index_of(q2));
      face(index_of(q2), index_of(p1), index_of(p2));
      if (mesh_type != FEM_MESH_OPEN) {
        const V3 k1 = projected(T[a1]), k2 = projected(T[a2]);
        add(k1, values[T[a1]]);
        add(k2, values[T[a2]]);
        face(index_of(q1), index_of(k1), index_of(q2));
        face(index_of(q2), index_of(k1), index_of(k2));
      }
    } else if (count == 3) {
      uint32_t c[3];
      for (int i = 0; i < 3; i++) {
        const V3 p = extruded(T[i]);
        add(p, values[T[i]]);
        c[i] = index_of(p);
      }
      face(c[0], c[1], c[2]);
      if (mesh_type != FEM_MESH_OPEN) {
        uint32_t pr[3];
        for (int i = 0; i < 3; i++) {
          const V3 np = projected(T[i]);
          add(np, threshold);
          pr[i] = index_of(np);
        }
        face(pr[0], pr[1], pr[2]);
      }
    }
  }
  std::string out;
  out.reserve(64 * pos.size() + 24 * (faces.size() / 3) + 256);
  out += "ply\nformat ascii 1.0\n";
  out += "element vertex " + std::to_string(pos.size()) + "\n";
  out += "property float x\nproperty float y\nproperty float z\n";
  out += "property float red\nproperty float green\nproperty float blue\n";
  out += "element face " + std::to_string(faces.size() / 3) + "\n";
  out += "property list uchar uint vertex_indices\nend_header\n";
  for (size_t i = 0; i < pos.size(); i++) {
    const V3 c = viridis(pval[i]);
    const float f6[6] = {pos[i].x, pos[i].y, pos[i].z, c.x, c.y, c.z};
    for (int k = 0; k < 6; k++) {
      if (k) out += ' ';
      put_float(out, f6[k]);
    }
    out += '\n';
  }
  for (size_t f = 0; f < faces.size(); f += 3)
    out += "3 " + std::to_string(faces[f]) + " " + std::to_string(faces[f + 1]) + " " + std::to_string(faces[f + 2]) +
           "\n";
  std::FILE* fp = std::fopen(path, "wb");
  if (!fp) {
    set_error(std::string("export_ply: cannot open ") + path + " for writing");
    return FEM_ERR_INVALID;
  }
  const size_t w = std::fwrite(out.data(), 1, out.size(), fp);
  const int cl = std::fclose(fp);
  if (w != out.size() || cl != 0) {
    set_error(std::string("export_ply: short write to ") + path);
    return FEM_ERR_INVALID;
  }
  return 0;
}

int surface_from_obj(const char* path, fem_surface** out) {
  fem_surface* s = new fem_surface();
  const int rc = read_obj(path, s);
  if (rc) {
    delete s;
    return rc;
  }
  *out = s;
  return 0;
}

void surface_free(fem_surface* s) { delete s; }

int surface_get(const fem_surface* s, fem_mesh* mesh, const float** normals, int32_t* num_boundary) {
  if (mesh) {
    mesh->num_vertices = (int32_t)(s->xyz.size() / 3);
    mesh->num_triangles = (int32_t)(s->tri.size() / 3);
    mesh->xyz = s->xyz.data();
    mesh->triangles = s->tri.data();
    mesh->on_boundary = s->onb.data();
  }
  if (normals) *normals = s->normals.data();
  if (num_boundary) *num_boundary = s->num_boundary;
  return 0;
}

}  // namespace femhip
