// Host-side mesh construction: the structured box generator and the
// node->element CSR builder.  Both must reproduce the reference bit for bit
// (coordinates, connectivity, CSR order), so they follow its definitions:
//   AddBoxLength  Domain_d.C:1136-1504
//   setNodElem    Domain_d.C:1508-1611
#include <cstddef>
#include <cstdint>
#include <vector>

namespace wfx {

// Structured 3D box.  Element counts truncate L/(2r) (Domain_d.C:1151-1166);
// node coordinates accumulate +2r per index step instead of i*2r
// (Domain_d.C:1215-1234), so the same rounding is reproduced.  Nodes run x
// fastest, then y, then z.  Hexa local order (Domain_d.C:1324-1332):
// nb1, nb1+1, nb2+1, nb2 on the lower layer, then the same +nnodz; six tets
// per cube take the hexa corners {0,1,3,5},{1,2,3,5},{0,5,3,4},{4,5,3,7},
// {5,6,3,7},{5,2,3,6} (Domain_d.C:1383-1388).
int box_length_mesh(const double V[3], const double L[3], double r, bool tritetra,
                    std::vector<double> &x, std::vector<int> &elnod, int &nodxelem) {
  const double step = 2.0 * r;
  const double fx = L[0] / step, fy = L[1] / step, fz = L[2] / step;
  // counts in int range before the conversion (also refuses NaN)
  if (!(fx >= 1.0 && fy >= 1.0 && fz >= 1.0 && fx < 2147483647.0 && fy < 2147483647.0 && fz < 2147483647.0))
    return -1;
  const int nx = (int)fx, ny = (int)fy, nz = (int)fz;
  const int px = nx + 1, py = ny + 1, pz = nz + 1;
  const int64_t nn = (int64_t)px * py * pz;
  const int64_t ncube = (int64_t)nx * ny * nz;
  const int64_t ne = tritetra ? 6 * ncube : ncube;
  if (nn > INT32_MAX || ne * (tritetra ? 4 : 8) > INT32_MAX) return -1;
  nodxelem = tritetra ? 4 : 8;
  x.resize((size_t)nn * 3);
  // one accumulated value per index along each axis, identical for every row
  std::vector<double> cx(px), cy(py), cz(pz);
  double acc = V[0];
  for (int i = 0; i < px; ++i) { cx[i] = acc; acc = acc + step; }
  acc = V[1];
  for (int j = 0; j < py; ++j) { cy[j] = acc; acc = acc + step; }
  acc = V[2];
  for (int k = 0; k < pz; ++k) { cz[k] = acc; acc = acc + 2 * r; }
  size_t p = 0;
  for (int k = 0; k < pz; ++k)
    for (int j = 0; j < py; ++j)
      for (int i = 0; i < px; ++i, ++p) {
        x[3 * p] = cx[i];
        x[3 * p + 1] = cy[j];
        x[3 * p + 2] = cz[k];
      }
  elnod.resize((size_t)ne * nodxelem);
  static const int kTet[6][4] = {{0, 1, 3, 5}, {1, 2, 3, 5}, {0, 5, 3, 4},
                                 {4, 5, 3, 7}, {5, 6, 3, 7}, {5, 2, 3, 6}};
  const int layer = px * py;
  size_t w = 0;
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        const int b1 = layer * k + px * j + i;
        const int b2 = layer * k + px * (j + 1) + i;
        const int c[8] = {b1, b1 + 1, b2 + 1, b2, b1 + layer, b1 + layer + 1, b2 + layer + 1, b2 + layer};
        if (!tritetra) {
          for (int a = 0; a < 8; ++a) elnod[w++] = c[a];
        } else {
          for (int t = 0; t < 6; ++t)
            for (int a = 0; a < 4; ++a) elnod[w++] = c[kTet[t][a]];
        }
      }
  return 0;
}

// node -> element CSR: every node lists the elements that reference it in
// ascending element order, with the local node index (Domain_d.C:1508-1611).
// A counting sort over (element, local) pairs in element order gives exactly
// that order.
int build_nodel(int N, int E, int n, const int *elnod, std::vector<int> &nodel,
                std::vector<int> &nodel_loc, std::vector<int> &offset, std::vector<int> &count) {
  count.assign(N, 0);
  const int64_t nnz = (int64_t)E * n;
  for (int64_t k = 0; k < nnz; ++k) {
    int g = elnod[k];
    if (g < 0 || g >= N) return -1;
    count[g]++;
  }
  offset.resize(N);
  int64_t tot = 0;
  for (int i = 0; i < N; ++i) {
    offset[i] = (int)tot;
    tot += count[i];
  }
  nodel.resize(tot);
  nodel_loc.resize(tot);
  std::vector<int> fill(N, 0);
  for (int e = 0; e < E; ++e)
    for (int a = 0; a < n; ++a) {
      int g = elnod[(int64_t)e * n + a];
      int pos = offset[g] + fill[g]++;
      nodel[pos] = e;
      nodel_loc[pos] = a;
    }
  return 0;
}

}  // namespace wfx
