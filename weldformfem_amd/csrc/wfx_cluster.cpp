// Host plan of the element clusters (see wfx_cluster.hpp).
#include "wfx_cluster.hpp"

#include <algorithm>
#include <cmath>
#include <numeric>
#include <thread>

namespace wfx {
namespace {

int find_root(std::vector<int> &parent, int i) {
  while (parent[i] != i) {
    parent[i] = parent[parent[i]];
    i = parent[i];
  }
  return i;
}

// 21 bits of v spread to every third bit
uint64_t spread3(uint64_t v) {
  v &= 0x1fffffULL;
  v = (v | (v << 32)) & 0x1f00000000ffffULL;
  v = (v | (v << 16)) & 0x1f0000ff0000ffULL;
  v = (v | (v << 8)) & 0x100f00f00f00f00fULL;
  v = (v | (v << 4)) & 0x10c30c30c30c30c3ULL;
  v = (v | (v << 2)) & 0x1249249249249249ULL;
  return v;
}

uint64_t quantize(double c, double lo, double h) {
  const double q = std::floor((c - lo) / h + 0.5);
  if (!(q > 0.0)) return 0;
  if (q > 262143.0) return 262143;
  return (uint64_t)q;
}

// Sort key of quantised coordinates q: bricks of 2^bits[0] x 2^bits[1] x
// 2^bits[2] cells along a Morton curve, cells within a brick x fastest (with
// no brick bits: the Morton code of the cell)
uint64_t brick_key(const uint64_t q[3], const int bits[3]) {
  uint64_t local = 0, b[3];
  int sh = 0;
  for (int d = 0; d < 3; ++d) {
    local |= (q[d] & ((1ULL << bits[d]) - 1)) << sh;
    sh += bits[d];
    b[d] = q[d] >> bits[d];
  }
  return ((spread3(b[0]) | (spread3(b[1]) << 1) | (spread3(b[2]) << 2)) << sh) | local;
}

// Slot tables of clusters [c0, c1): per cluster its sorted (node, t*8 + a)
// pairs give the slots (distinct nodes) and the entries grouped by slot.
// With write = false only the slot counts m[c] are taken.
void slot_tables(int n, const int *elnod, ClusterPlan &p, int c0, int c1, bool write, std::vector<int> &m) {
  std::vector<uint64_t> pairs;
  for (int c = c0; c < c1; ++c) {
    const int e0 = p.eoff[c], ne = p.eoff[c + 1] - e0;
    pairs.resize((size_t)ne * n);
    for (int t = 0; t < ne; ++t) {
      const int *en = elnod + (size_t)p.s2o[e0 + t] * n;
      for (int a = 0; a < n; ++a) pairs[(size_t)t * n + a] = ((uint64_t)en[a] << 16) | (uint64_t)(t * 8 + a);
    }
    std::sort(pairs.begin(), pairs.end());
    if (!write) {
      int cnt = 0;
      for (size_t k = 0; k < pairs.size(); ++k)
        if (k == 0 || (pairs[k] >> 16) != (pairs[k - 1] >> 16)) ++cnt;
      m[c] = cnt;
      continue;
    }
    uint16_t *lo = p.loff.data() + p.sbase[c] + c;
    uint16_t *en = p.ent.data() + (size_t)e0 * n;
    int *sn = p.snode.data() + p.sbase[c];
    int j = 0;
    for (size_t k = 0; k < pairs.size(); ++k) {
      if (k == 0 || (pairs[k] >> 16) != (pairs[k - 1] >> 16)) {
        sn[j] = (int)(pairs[k] >> 16);
        lo[j++] = (uint16_t)k;
      }
      const int q = (int)(pairs[k] & 0xffffULL);
      en[k] = (uint16_t)q;
      p.lslot[(size_t)(e0 + (q >> 3)) * n + (q & 7)] = (uint16_t)(j - 1);
    }
    lo[j] = (uint16_t)pairs.size();
  }
}

template <class F>
void parallel_ranges(int count, F &&f) {
  const int hw = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  const int nt = count < 64 ? 1 : hw;
  if (nt == 1) {
    f(0, count);
    return;
  }
  std::vector<std::thread> th;
  for (int k = 0; k < nt; ++k) {
    const int a = (int)((long)count * k / nt), b = (int)((long)count * (k + 1) / nt);
    th.emplace_back([&f, a, b] { f(a, b); });
  }
  for (auto &t : th) t.join();
}

}  // namespace

int plan_clusters(int N, int E, int n, const int *elnod, const double *x, int cs, ClusterPlan &p) {
  // connected components (elements sharing a node), ordered by their first element
  std::vector<int> parent(N);
  std::iota(parent.begin(), parent.end(), 0);
  for (int e = 0; e < E; ++e) {
    const int r0 = find_root(parent, elnod[(size_t)e * n]);
    for (int a = 1; a < n; ++a) {
      const int r = find_root(parent, elnod[(size_t)e * n + a]);
      if (r != r0) parent[r] = r0;
    }
  }
  std::vector<int> comp_of_root(N, -1), comp(E);
  int ncomp = 0;
  for (int e = 0; e < E; ++e) {
    const int r = find_root(parent, elnod[(size_t)e * n]);
    if (comp_of_root[r] < 0) comp_of_root[r] = ncomp++;
    comp[e] = comp_of_root[r];
  }
  std::vector<int> coff(ncomp + 1, 0);
  for (int e = 0; e < E; ++e) ++coff[comp[e] + 1];
  for (int k = 0; k < ncomp; ++k) coff[k + 1] += coff[k];
  std::vector<int> byc(E);
  {
    std::vector<int> fill(coff.begin(), coff.end() - 1);
    for (int e = 0; e < E; ++e) byc[fill[comp[e]]++] = e;  // ascending within each component
  }
  // element centroids and sizes (cube root of the node bounding box volume)
  std::vector<double> cen((size_t)3 * E);
  std::vector<double> size(E);
  for (int e = 0; e < E; ++e) {
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300}, s[3] = {0, 0, 0};
    for (int a = 0; a < n; ++a) {
      const double *q = x + 3 * (size_t)elnod[(size_t)e * n + a];
      for (int d = 0; d < 3; ++d) {
        s[d] += q[d];
        lo[d] = std::min(lo[d], q[d]);
        hi[d] = std::max(hi[d], q[d]);
      }
    }
    for (int d = 0; d < 3; ++d) cen[3 * (size_t)e + d] = s[d] / n;
    size[e] = std::cbrt(std::max(0.0, (hi[0] - lo[0]) * (hi[1] - lo[1]) * (hi[2] - lo[2])));
  }
  // each component along a Morton curve of bricks of its centroids,
  // quantised to its mean element size (a structured block maps to its
  // element indices, so a run of cs elements is one brick when the brick
  // holds cs cells), then cut into clusters
  p.s2o.resize(E);
  p.eoff.assign(1, 0);
  std::vector<std::pair<uint64_t, int>> key;
  for (int k = 0; k < ncomp; ++k) {
    const int a = coff[k], b = coff[k + 1];
    double lo[3] = {1e300, 1e300, 1e300}, hs = 0.0;
    for (int q = a; q < b; ++q) {
      const int e = byc[q];
      for (int d = 0; d < 3; ++d) lo[d] = std::min(lo[d], cen[3 * (size_t)e + d]);
      hs += size[e];
    }
    double h = hs / (b - a);
    if (!(h > 0.0) || !std::isfinite(h)) h = 1.0;
    key.resize(b - a);
    std::vector<uint64_t> cells(b - a);
    const int nobits[3] = {0, 0, 0};
    for (int q = a; q < b; ++q) {
      const int e = byc[q];
      const uint64_t qc[3] = {quantize(cen[3 * (size_t)e], lo[0], h), quantize(cen[3 * (size_t)e + 1], lo[1], h),
                              quantize(cen[3 * (size_t)e + 2], lo[2], h)};
      cells[q - a] = brick_key(qc, nobits);
    }
    // a brick holds about cs elements: 2^bits cells of the component's
    // occupancy (elements per occupied cell), half of the bits along x so
    // that a wavefront's elements run in long rows (their node gathers
    // coalesce), the rest split between y and z
    std::vector<uint64_t> uc(cells);
    std::sort(uc.begin(), uc.end());
    const double per_cell = (double)(b - a) / (double)(std::unique(uc.begin(), uc.end()) - uc.begin());
    const int tb = std::max(0, (int)std::lround(std::log2((double)cs / per_cell)));
    const int bits[3] = {(tb + 1) / 2, (tb / 2 + 1) / 2, (tb / 2) / 2};
    for (int q = a; q < b; ++q) {
      const int e = byc[q];
      const uint64_t qc[3] = {quantize(cen[3 * (size_t)e], lo[0], h), quantize(cen[3 * (size_t)e + 1], lo[1], h),
                              quantize(cen[3 * (size_t)e + 2], lo[2], h)};
      key[q - a] = {brick_key(qc, bits), e};
    }
    std::sort(key.begin(), key.end());
    for (int q = a; q < b; q += cs) {
      const int r = std::min(b, q + cs);
      for (int t = q; t < r; ++t) p.s2o[t] = key[t - a].second;
      std::sort(p.s2o.begin() + q, p.s2o.begin() + r);  // ascending original order within a cluster
      p.eoff.push_back(r);
    }
  }
  p.o2s.resize(E);
  for (int s = 0; s < E; ++s) p.o2s[p.s2o[s]] = s;
  // slot counts; a cluster touching more than slot_cap(n) nodes is halved
  // until none does (a single element has n <= slot_cap(n) nodes)
  std::vector<int> m;
  for (;;) {
    const int C = p.clusters();
    m.assign(C, 0);
    parallel_ranges(C, [&](int c0, int c1) { slot_tables(n, elnod, p, c0, c1, false, m); });
    std::vector<int> eoff(1, 0);
    bool split = false;
    for (int c = 0; c < C; ++c) {
      if (m[c] > slot_cap(n)) {
        eoff.push_back(p.eoff[c] + (p.eoff[c + 1] - p.eoff[c]) / 2);
        split = true;
      }
      eoff.push_back(p.eoff[c + 1]);
    }
    if (!split) break;
    p.eoff.swap(eoff);
  }
  const int C = p.clusters();
  p.max_slots = C ? *std::max_element(m.begin(), m.end()) : 0;
  p.sbase.assign(C + 1, 0);
  for (int c = 0; c < C; ++c) p.sbase[c + 1] = p.sbase[c] + m[c];
  const long S = p.slots();
  p.loff.assign((size_t)S + C, 0);
  p.ent.assign((size_t)E * n, 0);
  p.snode.assign((size_t)S, 0);
  p.lslot.assign((size_t)E * n, 0);
  parallel_ranges(C, [&](int c0, int c1) { slot_tables(n, elnod, p, c0, c1, true, m); });
  // node -> slots, in cluster order
  p.nsl_off.assign((size_t)N + 1, 0);
  for (int c = 0; c < C; ++c) {
    const int e0 = p.eoff[c];
    const uint16_t *lo = p.loff.data() + p.sbase[c] + c;
    const uint16_t *en = p.ent.data() + (size_t)e0 * n;
    for (int j = 0; j < m[c]; ++j) {
      const int q = en[lo[j]];
      ++p.nsl_off[(size_t)elnod[(size_t)p.s2o[e0 + (q >> 3)] * n + (q & 7)] + 1];
    }
  }
  for (int i = 0; i < N; ++i) p.nsl_off[i + 1] += p.nsl_off[i];
  p.nsl.assign((size_t)S, 0);
  std::vector<int> fill(p.nsl_off.begin(), p.nsl_off.end() - 1);
  for (int c = 0; c < C; ++c) {
    const int e0 = p.eoff[c];
    const uint16_t *lo = p.loff.data() + p.sbase[c] + c;
    const uint16_t *en = p.ent.data() + (size_t)e0 * n;
    for (int j = 0; j < m[c]; ++j) {
      const int q = en[lo[j]];
      const int node = elnod[(size_t)p.s2o[e0 + (q >> 3)] * n + (q & 7)];
      p.nsl[fill[node]++] = p.sbase[c] + j;
    }
  }
  p.spos.assign((size_t)S, 0);
  for (long k = 0; k < S; ++k) p.spos[p.nsl[k]] = (int)k;
  return 0;
}

}  // namespace wfx
