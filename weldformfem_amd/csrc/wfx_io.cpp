// The on-disk formats either side of the step: LS-DYNA keyword meshes in
// (*NODE, *ELEMENT_SOLID; what CreateFromLSDyna consumes, Domain_d.C:1647-1699)
// and legacy VTK unstructured-grid frames out (the fields VTKWriter.C:365-625
// writes that this library keeps).  Host code only.
#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <sstream>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/wfx.h"
#include "wfx_vtk.hpp"

namespace wfx {
void set_error(const std::string &msg);
}

namespace {

// A card field as a character range of the file buffer (no copies).
struct Field {
  const char *b, *e;
};

inline void trim(const char *&b, const char *&e) {
  while (b < e && std::isspace((unsigned char)*b)) ++b;
  while (e > b && std::isspace((unsigned char)e[-1])) --e;
}

// Card fields of the line [b, e): comma separated (free format) or fixed
// columns of the given widths; a fixed line too short for its widths ends
// early.  Returns the number of fields.
int split_fields(const char *b, const char *e, const int *widths, int nw, Field *out) {
  int n = 0;
  if (std::memchr(b, ',', (size_t)(e - b))) {
    const char *p = b;
    while (n < nw) {
      const char *q = (const char *)std::memchr(p, ',', (size_t)(e - p));
      if (!q) q = e;
      Field f{p, q};
      trim(f.b, f.e);
      out[n++] = f;
      if (q == e) break;
      p = q + 1;
    }
    return n;
  }
  const char *p = b;
  for (int k = 0; k < nw && p < e; ++k) {
    const char *q = (e - p > widths[k]) ? p + widths[k] : e;
    Field f{p, q};
    trim(f.b, f.e);
    out[n++] = f;
    p = q;
  }
  return n;
}

// the whole (trimmed) field must be the number: an optional sign and at most
// 18 decimal digits
bool to_long(Field f, long &v) {
  const char *p = f.b;
  bool neg = false;
  if (p < f.e && (*p == '-' || *p == '+')) neg = *p++ == '-';
  if (p == f.e || f.e - p > 18) return false;
  long r = 0;
  for (; p < f.e; ++p) {
    const unsigned d = (unsigned)(*p - '0');
    if (d > 9) return false;
    r = 10 * r + (long)d;
  }
  v = neg ? -r : r;
  return true;
}
bool to_double(Field f, double &v) {
  const size_t len = (size_t)(f.e - f.b);
  if (len == 0 || len > 63) return false;
  char buf[64];
  std::memcpy(buf, f.b, len);
  buf[len] = '\0';
  char *end = nullptr;
  v = std::strtod(buf, &end);
  return *end == '\0';
}

struct KMesh {
  std::vector<long> node_id;
  std::vector<double> x;
  std::vector<long> enodes;   // node ids of all elements, element after element
  std::vector<long> estart;   // element k: enodes[estart[k] .. estart[k + 1])
  size_t elem_count() const { return estart.empty() ? 0 : estart.size() - 1; }
};

bool read_file(const char *path, std::vector<char> &buf) {
  FILE *f = std::fopen(path, "rb");
  if (!f) return false;
  std::fseek(f, 0, SEEK_END);
  const long size = std::ftell(f);
  std::fseek(f, 0, SEEK_SET);
  buf.resize(size > 0 ? (size_t)size : 0);
  const size_t got = size > 0 ? std::fread(buf.data(), 1, (size_t)size, f) : 0;
  std::fclose(f);
  return got == buf.size();
}

// Reads *NODE (nid, x, y, z: I8, 3 x E16) and *ELEMENT_SOLID (eid, pid,
// n1..n8: I8 each, or the two-line form with the nodes on the second card).
// Cards that are neither fixed nor comma separated are read as blank
// separated numbers.
int parse_k(const char *path, KMesh &m, std::string &err) {
  std::vector<char> buf;
  if (!read_file(path, buf)) {
    err = std::string("cannot open ") + path;
    return WFX_ERR_ARG;
  }
  static const int kNodeW[6] = {8, 16, 16, 16, 8, 8};
  static const int kSolidW[10] = {8, 8, 8, 8, 8, 8, 8, 8, 8, 8};
  enum { NONE, NODE, SOLID } sec = NONE;
  long lineno = 0;
  bool pending = false;  // element card whose nodes are on the next line
  m.estart.assign(1, 0);
  const char *p = buf.data(), *end = p + buf.size();
  Field f[10];
  while (p < end) {
    const char *nl = (const char *)std::memchr(p, '\n', (size_t)(end - p));
    const char *b = p, *e = nl ? nl : end;
    p = nl ? nl + 1 : end;
    ++lineno;
    if (e > b && e[-1] == '\r') --e;
    const char *t = b;  // blank lines (spaces only) are skipped like empty ones
    while (t < e && std::isspace((unsigned char)*t)) ++t;
    if (t == e || *b == '$') continue;
    if (*b == '*') {
      const char *kb = b, *ke = e;
      trim(kb, ke);
      std::string kw(kb, ke);
      for (auto &ch : kw) ch = (char)std::toupper((unsigned char)ch);
      sec = kw == "*NODE" ? NODE : (kw == "*ELEMENT_SOLID" ? SOLID : NONE);
      pending = false;
      continue;
    }
    if (sec == NODE) {
      long id;
      double c[3];
      const int nf = split_fields(b, e, kNodeW, 6, f);
      if (nf < 4 || !to_long(f[0], id) || !to_double(f[1], c[0]) || !to_double(f[2], c[1]) ||
          !to_double(f[3], c[2])) {
        std::stringstream ss(std::string(b, e));  // blank separated
        if (!(ss >> id >> c[0] >> c[1] >> c[2])) {
          err = "bad *NODE card at line " + std::to_string(lineno);
          return WFX_ERR_ARG;
        }
      }
      m.node_id.push_back(id);
      m.x.insert(m.x.end(), c, c + 3);
    } else if (sec == SOLID) {
      long v[16];
      int nv = 0;
      bool ok = true;
      const int nf = split_fields(b, e, kSolidW, 10, f);
      for (int k = 0; k < nf && ok; ++k) {
        if (f[k].b == f[k].e) continue;
        ok = to_long(f[k], v[nv]);
        if (ok) ++nv;
      }
      if (!ok) {  // blank separated
        nv = 0;
        std::stringstream ss(std::string(b, e));
        long q;
        while (nv < 16 && ss >> q) v[nv++] = q;
        if (!ss.eof()) {
          err = "bad *ELEMENT_SOLID card at line " + std::to_string(lineno);
          return WFX_ERR_ARG;
        }
      }
      if (pending) {  // second card of the two-line form
        m.enodes.insert(m.enodes.end(), v, v + nv);
        m.estart.push_back((long)m.enodes.size());
        pending = false;
      } else if (nv == 2) {
        pending = true;
      } else if (nv >= 6) {
        m.enodes.insert(m.enodes.end(), v + 2, v + nv);
        m.estart.push_back((long)m.enodes.size());
      } else {
        err = "bad *ELEMENT_SOLID card at line " + std::to_string(lineno);
        return WFX_ERR_ARG;
      }
    }
  }
  if (m.node_id.empty() || m.elem_count() == 0) {
    err = std::string("no *NODE or *ELEMENT_SOLID data in ") + path;
    return WFX_ERR_ARG;
  }
  return WFX_OK;
}

// Node ids become indices in their order of appearance under *NODE (a direct
// table when the ids are dense, a hash map otherwise).  A solid whose nodes
// 4..8 repeat node 4 (the keyword form of a tetrahedron) keeps its first four
// nodes.  All elements must be of one kind.
int to_arrays(const KMesh &m, int &nodxelem, std::vector<int> &elnod, std::string &err) {
  const size_t N = m.node_id.size();
  long lo = m.node_id[0], hi = m.node_id[0];
  for (long id : m.node_id) {
    lo = id < lo ? id : lo;
    hi = id > hi ? id : hi;
  }
  const bool dense = hi - lo < 4L * (long)N + 64;
  std::vector<int> table;
  std::unordered_map<long, int> idx;
  if (dense) table.assign((size_t)(hi - lo + 1), -1);
  else idx.reserve(N * 2);
  for (size_t k = 0; k < N; ++k) {
    const long id = m.node_id[k];
    bool fresh;
    if (dense) {
      fresh = table[(size_t)(id - lo)] < 0;
      if (fresh) table[(size_t)(id - lo)] = (int)k;
    } else {
      fresh = idx.emplace(id, (int)k).second;
    }
    if (!fresh) {
      err = "duplicate node id " + std::to_string(id);
      return WFX_ERR_ARG;
    }
  }
  auto find = [&](long id) -> int {
    if (dense) return (id < lo || id > hi) ? -1 : table[(size_t)(id - lo)];
    auto it = idx.find(id);
    return it == idx.end() ? -1 : it->second;
  };
  nodxelem = 0;
  elnod.clear();
  const size_t E = m.elem_count();
  for (size_t e = 0; e < E; ++e) {
    const long *nd = m.enodes.data() + m.estart[e];
    const long cnt = m.estart[e + 1] - m.estart[e];
    int n;
    if (cnt == 4) n = 4;
    else if (cnt == 8) {
      bool tet = true;
      for (int k = 4; k < 8; ++k) tet = tet && nd[k] == nd[3];
      n = tet ? 4 : 8;
    } else {
      err = "element " + std::to_string(e + 1) + " has " + std::to_string(cnt) + " nodes";
      return WFX_ERR_ARG;
    }
    if (nodxelem == 0) {
      nodxelem = n;
      elnod.reserve(E * (size_t)n);
    }
    if (n != nodxelem) {
      err = "mixed hexahedra and tetrahedra are not supported";
      return WFX_ERR_ARG;
    }
    for (int k = 0; k < n; ++k) {
      const int q = find(nd[k]);
      if (q < 0) {
        err = "element " + std::to_string(e + 1) + " refers to unknown node " + std::to_string(nd[k]);
        return WFX_ERR_ARG;
      }
      elnod.push_back(q);
    }
  }
  return WFX_OK;
}

// parse + index conversion, one pass over the file
int load_k(const char *path, KMesh &m, int &n, std::vector<int> &el) {
  std::string err;
  int rc = parse_k(path, m, err);
  if (rc == WFX_OK) rc = to_arrays(m, n, el, err);
  if (rc != WFX_OK) wfx::set_error(err);
  return rc;
}

}  // namespace

namespace wfx {

namespace {

// Threads that format a frame: half the host's hardware threads, at most 8
// (the step loop keeps a core for its launches).
int writer_threads() {
  const unsigned h = std::thread::hardware_concurrency();
  const int t = (int)(h / 2);
  return t < 1 ? 1 : (t > 8 ? 8 : t);
}

// Items [0, count) formatted by item(i, out) (appending to out) and written
// in order: blocks of 2^18 items, each block split over the writer threads.
template <class Item>
bool write_items(FILE *f, long count, Item item) {
  constexpr long kBlock = 1L << 18;
  const int T = writer_threads();
  std::vector<std::string> part(T);
  for (long b0 = 0; b0 < count; b0 += kBlock) {
    const long b1 = b0 + kBlock < count ? b0 + kBlock : count;
    auto run = [&](int t) {
      const long lo = b0 + (b1 - b0) * t / T, hi = b0 + (b1 - b0) * (t + 1) / T;
      part[t].clear();
      for (long i = lo; i < hi; ++i) item(i, part[t]);
    };
    if (T == 1 || b1 - b0 < 4096) {  // small: one thread
      part[0].clear();
      for (long i = b0; i < b1; ++i) item(i, part[0]);
      if (std::fwrite(part[0].data(), 1, part[0].size(), f) != part[0].size()) return false;
      continue;
    }
    std::vector<std::thread> th;
    for (int t = 1; t < T; ++t) th.emplace_back(run, t);
    run(0);
    for (auto &x : th) x.join();
    for (auto &q : part)
      if (std::fwrite(q.data(), 1, q.size(), f) != q.size()) return false;
  }
  return true;
}

inline void put3(std::string &out, double a, double b, double c) {
  char buf[96];
  const int k = std::snprintf(buf, sizeof(buf), "%.9e %.9e %.9e\n", a, b, c);
  out.append(buf, (size_t)k);
}

}  // namespace

// VTKWriter::writeFile's legacy unstructured grid (VTKWriter.C:365-625),
// restricted to the fields this library keeps; the sections are formatted by
// several threads and written in order (the same bytes as one thread).
int write_vtk_frame(const char *path, const VtkFrame &fr, std::string &err) {
  const int N = fr.N, E = fr.E, n = fr.n;
  FILE *f = std::fopen(path, "w");
  if (!f) {
    err = std::string("cannot write ") + path;
    return 1;
  }
  bool ok = true;
  std::fprintf(f, "# vtk DataFile Version 3.0\nwfx explicit step\nASCII\nDATASET UNSTRUCTURED_GRID\n");
  std::fprintf(f, "POINTS %d double\n", N);
  ok = ok && write_items(f, N, [&](long i, std::string &o) { put3(o, fr.x[3 * i], fr.x[3 * i + 1], fr.x[3 * i + 2]); });
  std::fprintf(f, "CELLS %d %ld\n", E, (long)E * (n + 1));
  ok = ok && write_items(f, E, [&](long e, std::string &o) {
    char buf[128];
    int k = std::snprintf(buf, sizeof(buf), "%d", n);
    for (int a = 0; a < n; ++a) k += std::snprintf(buf + k, sizeof(buf) - (size_t)k, " %d", fr.elnod[e * n + a]);
    buf[k++] = '\n';
    o.append(buf, (size_t)k);
  });
  std::fprintf(f, "CELL_TYPES %d\n", E);
  ok = ok && write_items(f, E, [&](long, std::string &o) { o.append(n == 8 ? "12\n" : "10\n"); });
  auto vec = [&](const char *name, const std::vector<double> &q) {
    std::fprintf(f, "VECTORS %s double\n", name);
    ok = ok && write_items(f, N, [&](long i, std::string &o) { put3(o, q[3 * i], q[3 * i + 1], q[3 * i + 2]); });
  };
  auto scal = [&](const char *name, const std::vector<double> &q, long count) {
    std::fprintf(f, "SCALARS %s double 1\nLOOKUP_TABLE default\n", name);
    ok = ok && write_items(f, count, [&](long i, std::string &o) {
      char buf[40];
      const int k = std::snprintf(buf, sizeof(buf), "%.9e\n", q[i]);
      o.append(buf, (size_t)k);
    });
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
  ok = ok && write_items(f, E, [&](long e, std::string &o) {
    const double *q = &fr.sigma[6L * e];  // xx yy zz xy yz xz
    put3(o, q[0], q[3], q[5]);
    put3(o, q[3], q[1], q[4]);
    put3(o, q[5], q[4], q[2]);
  });
  ok = ok && std::ferror(f) == 0;
  ok = (std::fclose(f) == 0) && ok;
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
  std::vector<int> el;
  int n = 0;
  int rc = load_k(path, m, n, el);
  if (rc != WFX_OK) return rc;
  *node_count = (int)m.node_id.size();
  *elem_count = (int)m.elem_count();
  *nodxelem = n;
  if (x) std::memcpy(x, m.x.data(), m.x.size() * sizeof(double));
  if (elnod) std::memcpy(elnod, el.data(), el.size() * sizeof(int));
  return WFX_OK;
}

int wfx_read_lsdyna_k(wfx_ctx *ctx, const char *path) {
  if (!path) {
    wfx::set_error("null path");
    return WFX_ERR_ARG;
  }
  KMesh m;
  std::vector<int> el;
  int n = 0;
  int rc = load_k(path, m, n, el);
  if (rc != WFX_OK) return rc;
  return wfx_set_mesh(ctx, (int)m.node_id.size(), (int)m.elem_count(), n, m.x.data(), el.data());
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
