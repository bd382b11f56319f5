// Host driver of the element-cluster planner (weldformfem_amd/csrc/wfx_cluster.cpp)
// for the CPU tests: reads a mesh, writes the plan's arrays as raw int32 files.
//   cluster_plan_check N E n cs elnod.i32 x.f64 out_prefix
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "../../weldformfem_amd/csrc/wfx_cluster.hpp"

template <class T>
static bool read_all(const char *path, std::vector<T> &v, size_t count) {
  FILE *f = std::fopen(path, "rb");
  if (!f) return false;
  v.resize(count);
  const size_t got = std::fread(v.data(), sizeof(T), count, f);
  std::fclose(f);
  return got == count;
}

template <class T>
static bool write_all(const std::string &path, const std::vector<T> &v) {
  FILE *f = std::fopen(path.c_str(), "wb");
  if (!f) return false;
  std::vector<int> w(v.begin(), v.end());
  const size_t put = std::fwrite(w.data(), sizeof(int), w.size(), f);
  std::fclose(f);
  return put == w.size();
}

int main(int argc, char **argv) {
  if (argc != 8) {
    std::fprintf(stderr, "usage: %s N E n cs elnod.i32 x.f64 out_prefix\n", argv[0]);
    return 2;
  }
  const int N = std::atoi(argv[1]), E = std::atoi(argv[2]), n = std::atoi(argv[3]), cs = std::atoi(argv[4]);
  std::vector<int> elnod;
  std::vector<double> x;
  if (!read_all(argv[5], elnod, (size_t)E * n) || !read_all(argv[6], x, (size_t)N * 3)) {
    std::fprintf(stderr, "cannot read the mesh\n");
    return 2;
  }
  wfx::ClusterPlan p;
  if (wfx::plan_clusters(N, E, n, elnod.data(), x.data(), cs, p)) return 1;
  const std::string o = argv[7];
  bool ok = write_all(o + "s2o", p.s2o) && write_all(o + "eoff", p.eoff) && write_all(o + "sbase", p.sbase) &&
            write_all(o + "loff", p.loff) && write_all(o + "ent", p.ent) && write_all(o + "nsl_off", p.nsl_off) &&
            write_all(o + "nsl", p.nsl) && write_all(o + "snode", p.snode) && write_all(o + "lslot", p.lslot) &&
            write_all(o + "spos", p.spos);
  std::printf("%d %ld %d %d\n", p.clusters(), p.slots(), p.max_slots, wfx::slot_cap(n));
  return ok ? 0 : 1;
}
