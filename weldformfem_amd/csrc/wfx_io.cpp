// The on-disk formats either side of the step: LS-DYNA keyword meshes in
// (*NODE, *ELEMENT_SOLID; what CreateFromLSDyna consumes, Domain_d.C:1647-1699)
// and legacy VTK unstructured-grid frames out (the fields VTKWriter.C:365-625
// writes that this build keeps).  Host code only.
#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/wfx.h"
#include "wfx_vtk.hpp"

namespace wfx {
void set_error(const std::string &msg);
}

namespace {

std::string trim(const std::string &s) {
  size_t a = 0, b = s.size();
  while (a < b && std::isspace((unsigned char)s[a])) ++a;
  while (b > a && std::isspace((unsigned char)s[b - 1])) --b;
  return s.substr(a, b - a);
}

// Card fields: comma separated (free format) or fixed columns of the given
// widths; a fixed line too short for its widths falls back to blanks.
std::vector<std::string> fields(const std::string &line, const std::vector<int> &widths) {
  std::vector<std::string> out;
  if (line.find(',') != std::string::npos) {
    std::stringstream ss(line);
    std::string f;
    while (std::getline(ss, f, ',')) out.push_back(trim(f));
    return out;
  }
  size_t pos = 0;
  for (int w : widths) {
    if (pos >= line.size()) break;
    out.push_back(trim(line.substr(pos, w)));
    pos += w;
  }
  return out;
}

bool to_long(const std::string &s, long &v) {
  if (s.empty()) return false;
  char *end = nullptr;
  v = std::strtol(s.c_str(), &end, 10);
  return end && *end == '\0';
}
bool to_double(const std::string &s, double &v) {
  if (s.empty()) return false;
  char *end = nullptr;
  v = std::strtod(s.c_str(), &end);
  return end && *end == '\0';
}

struct KMesh {
  std::vector<long> node_id;
  std::vector<double> x;
  std::vector<std::vector<long>> elem;  // node ids per element
};

// Reads *NODE (nid, x, y, z: I8, 3 x E16) and *ELEMENT_SOLID (eid, pid,
// n1..n8: I8 each, or the two-line form with the nodes on the second card).
int parse_k(const char *path, KMesh &m, std::string &err) {
  std::ifstream in(path);
  if (!in) {
    err = std::string("cannot open ") + path;
    return WFX_ERR_ARG;
  }
  enum { NONE, NODE, SOLID } sec = NONE;
  std::string line;
  long lineno = 0;
  std::vector<long> pending;  // element card whose nodes are on the next line
  while (std::getline(in, line)) {
    ++lineno;
    if (!line.empty() && line.back() == '\r') line.pop_back();
    if (line.empty() || line[0] == '$') continue;
    if (line[0] == '*') {
      std::string kw = trim(line);
      for (auto &ch : kw) ch = (char)std::toupper((unsigned char)ch);
      sec = kw == "*NODE" ? NODE : (kw == "*ELEMENT_SOLID" ? SOLID : NONE);
      pending.clear();
      continue;
    }
    if (sec == NODE) {
      auto f = fields(line, {8, 16, 16, 16, 8, 8});
      long id;
      double c[3];
      if (f.size() < 4 || !to_long(f[0], id) || !to_double(f[1], c[0]) || !to_double(f[2], c[1]) ||
          !to_double(f[3], c[2])) {
        // free format separated by blanks
        std::stringstream ss(line);
        if (!(ss >> id >> c[0] >> c[1] >> c[2])) {
          err = "bad *NODE card at line " + std::to_string(lineno);
          return WFX_ERR_ARG;
        }
      }
      m.node_id.push_back(id);
      m.x.insert(m.x.end(), c, c + 3);
    } else if (sec == SOLID) {
      auto f = fields(line, {8, 8, 8, 8, 8, 8, 8, 8, 8, 8});
      std::vector<long> v;
      bool ok = true;
      for (auto &s : f) {
        long q;
        if (s.empty()) continue;
        if (!to_long(s, q)) { ok = false; break; }
        v.push_back(q);
      }
      if (!ok) {  // free format separated by blanks
        v.clear();
        std::stringstream ss(line);
        long q;
        while (ss >> q) v.push_back(q);
        if (!ss.eof()) {
          err = "bad *ELEMENT_SOLID card at line " + std::to_string(lineno);
          return WFX_ERR_ARG;
        }
      }
      if (!pending.empty()) {  // second card of the two-line form
        std::vector<long> nodes(v.begin(), v.end());
        m.elem.push_back(nodes);
        pending.clear();
      } else if (v.size() == 2) {
        pending = v;
      } else if (v.size() >= 6) {
        m.elem.emplace_back(v.begin() + 2, v.end());
      } else {
        err = "bad *ELEMENT_SOLID card at line " + std::to_string(lineno);
        return WFX_ERR_ARG;
      }
    }
  }
  if (m.node_id.empty() || m.elem.empty()) {
    err = std::string("no *NODE or *ELEMENT_SOLID data in ") + path;
    return WFX_ERR_ARG;
  }
  return WFX_OK;
}

// Node ids become indices in their order of appearance under *NODE.  A solid
// whose nodes 4..8 repeat node 4 (the keyword form of a tetrahedron) keeps its
// first four nodes.  All elements must be of one kind.
int to_arrays(const KMesh &m, int &nodxelem, std::vector<int> &elnod, std::string &err) {
  std::unordered_map<long, int> idx;
  idx.reserve(m.node_id.size() * 2);
  for (size_t k = 0; k < m.node_id.size(); ++k)
    if (!idx.emplace(m.node_id[k], (int)k).second) {
      err = "duplicate node id " + std::to_string(m.node_id[k]);
      return WFX_ERR_ARG;
    }
  nodxelem = 0;
  elnod.clear();
  for (size_t e = 0; e < m.elem.size(); ++e) {
    const auto &nd = m.elem[e];
    int n;
    if (nd.size() == 4) n = 4;
    else if (nd.size() == 8) {
      bool tet = true;
      for (int k = 4; k < 8; ++k) tet = tet && nd[k] == nd[3];
      n = tet ? 4 : 8;
    } else {
      err = "element " + std::to_string(e + 1) + " has " + std::to_string(nd.size()) + " nodes";
      return WFX_ERR_ARG;
    }
    if (nodxelem == 0) nodxelem = n;
    if (n != nodxelem) {
      err = "mixed hexahedra and tetrahedra are not supported";
      return WFX_ERR_ARG;
    }
    for (int k = 0; k < n; ++k) {
      auto it = idx.find(nd[k]);
      if (it == idx.end()) {
        err = "element " + std::to_string(e + 1) + " refers to unknown node " + std::to_string(nd[k]);
        return WFX_ERR_ARG;
      }
      elnod.push_back(it->second);
    }
  }
  return WFX_OK;
}

}  // namespace

namespace wfx {

// VTKWriter::writeFile's legacy unstructured grid (VTKWriter.C:365-625),
// restricted to the fields this library keeps
int write_vtk_frame(const char *path, const VtkFrame &fr, std::string &err) {
  const int N = fr.N, E = fr.E, n = fr.n;
  FILE *f = std::fopen(path, "w");
  if (!f) {
    err = std::string("cannot write ") + path;
    return 1;
  }
  std::fprintf(f, "# vtk DataFile Version 3.0\nwfx explicit step\nASCII\nDATASET UNSTRUCTURED_GRID\n");
  std::fprintf(f, "POINTS %d double\n", N);
  for (int i = 0; i < N; ++i) std::fprintf(f, "%.9e %.9e %.9e\n", fr.x[3 * i], fr.x[3 * i + 1], fr.x[3 * i + 2]);
  std::fprintf(f, "CELLS %d %ld\n", E, (long)E * (n + 1));
  for (int e = 0; e < E; ++e) {
    std::fprintf(f, "%d", n);
    for (int k = 0; k < n; ++k) std::fprintf(f, " %d", fr.elnod[(long)e * n + k]);
    std::fprintf(f, "\n");
  }
  std::fprintf(f, "CELL_TYPES %d\n", E);
  for (int e = 0; e < E; ++e) std::fprintf(f, "%d\n", n == 8 ? 12 : 10);
  auto vec = [&](const char *name, const std::vector<double> &q) {
    std::fprintf(f, "VECTORS %s double\n", name);
    for (int i = 0; i < N; ++i) std::fprintf(f, "%.9e %.9e %.9e\n", q[3 * i], q[3 * i + 1], q[3 * i + 2]);
  };
  auto scal = [&](const char *name, const std::vector<double> &q, int count) {
    std::fprintf(f, "SCALARS %s double 1\nLOOKUP_TABLE default\n", name);
    for (int i = 0; i < count; ++i) std::fprintf(f, "%.9e\n", q[i]);
  };
  std::fprintf(f, "POINT_DATA %d\n", N);
  vec("DISP", fr.u);
  vec("Velocity", fr.v);
  vec("Acceleration", fr.a);
  scal("nod_mass", fr.mdiag, N);
  std::fprintf(f, "CELL_DATA %d\n", E);
  scal("pressure", fr.p, E);
  scal("pl_strain", fr.pl_strain, E);
  scal("sigma_y", fr.sigma_y, E);
  scal("Vol", fr.vol, E);
  std::fprintf(f, "TENSORS SIGMA double\n");
  for (int e = 0; e < E; ++e) {
    const double *q = &fr.sigma[6L * e];  // xx yy zz xy yz xz
    std::fprintf(f, "%.9e %.9e %.9e\n%.9e %.9e %.9e\n%.9e %.9e %.9e\n", q[0], q[3], q[5], q[3], q[1], q[4], q[5],
                 q[4], q[2]);
  }
  const bool ok = std::ferror(f) == 0;
  std::fclose(f);
  if (!ok) {
    err = std::string("write error on ") + path;
    return 1;
  }
  return 0;
}

}  // namespace wfx

extern "C" {

int wfx_parse_lsdyna_k(const char *path, int *node_count, int *elem_count, int *nodxelem, double *x, int *elnod) {
  if (!path || !node_count || !elem_count || !nodxelem) {
    wfx::set_error("null argument");
    return WFX_ERR_ARG;
  }
  KMesh m;
  std::string err;
  int rc = parse_k(path, m, err);
  std::vector<int> el;
  int n = 0;
  if (rc == WFX_OK) rc = to_arrays(m, n, el, err);
  if (rc != WFX_OK) {
    wfx::set_error(err);
    return rc;
  }
  *node_count = (int)m.node_id.size();
  *elem_count = (int)m.elem.size();
  *nodxelem = n;
  if (x) std::memcpy(x, m.x.data(), m.x.size() * sizeof(double));
  if (elnod) std::memcpy(elnod, el.data(), el.size() * sizeof(int));
  return WFX_OK;
}

int wfx_read_lsdyna_k(wfx_ctx *ctx, const char *path) {
  int N = 0, E = 0, n = 0;
  int rc = wfx_parse_lsdyna_k(path, &N, &E, &n, nullptr, nullptr);
  if (rc != WFX_OK) return rc;
  std::vector<double> x(3L * N);
  std::vector<int> el((long)E * n);
  rc = wfx_parse_lsdyna_k(path, &N, &E, &n, x.data(), el.data());
  if (rc != WFX_OK) return rc;
  return wfx_set_mesh(ctx, N, E, n, x.data(), el.data());
}

// Legacy ASCII VTK frame: current positions, connectivity (VTK_HEXAHEDRON 12,
// VTK_TETRA 10); point data DISP, Velocity, Acceleration, nod_mass; cell data
// pressure, pl_strain, sigma_y, Vol and the stress tensor.
int wfx_write_vtk(wfx_ctx *ctx, const char *path, wfx_stream s) {
  (void)s;
  wfx::VtkFrame fr;
  int rc = wfx_get_sizes(ctx, &fr.N, &fr.E, &fr.n);
  if (rc != WFX_OK) return rc;
  if (!path) {
    wfx::set_error("null path");
    return WFX_ERR_ARG;
  }
  auto get = [&](const char *name, std::vector<double> &out) {
    long cnt = 0;
    int is_int = 0;
    int r = wfx_field_count(ctx, name, &cnt, &is_int);
    if (r != WFX_OK) return r;
    out.resize(cnt);
    return wfx_get_field(ctx, name, out.data(), cnt);
  };
  for (auto pr : {std::make_pair("x", &fr.x), {"u", &fr.u}, {"v", &fr.v}, {"a", &fr.a}, {"mdiag", &fr.mdiag},
                  {"p", &fr.p}, {"pl_strain", &fr.pl_strain}, {"sigma_y", &fr.sigma_y}, {"vol", &fr.vol},
                  {"sigma", &fr.sigma}}) {
    rc = get(pr.first, *pr.second);
    if (rc != WFX_OK) return rc;
  }
  fr.elnod.resize((size_t)fr.E * fr.n);
  rc = wfx_get_field(ctx, "elnod", fr.elnod.data(), (long)fr.elnod.size());
  if (rc != WFX_OK) return rc;
  std::string err;
  if (wfx::write_vtk_frame(path, fr, err)) {
    wfx::set_error(err);
    return WFX_ERR_ARG;
  }
  return WFX_OK;
}

}  // extern "C"
