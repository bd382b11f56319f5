// C-ABI of the wfx library: context lifetime, mesh and set-up, the per-step
// Domain_d API and state transfer.  Kernels live in wfx_k64.hip / wfx_k32.hip.
#include <algorithm>
#include <array>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/wfx.h"
#include "wfx_cluster.hpp"
#include "wfx_internal.hpp"
#include "wfx_launch.hpp"

namespace wfx {

static thread_local std::string g_err;
void set_error(const std::string &msg) { g_err = msg; }

int box_length_mesh(const double V[3], const double L[3], double r, bool tritetra,
                    std::vector<double> &x, std::vector<int> &elnod, int &nodxelem);
int build_nodel(int N, int E, int n, const int *elnod, std::vector<int> &nodel,
                std::vector<int> &nodel_loc, std::vector<int> &offset, std::vector<int> &count);

// frees one of the context's allocations (before it grows or is replaced)
template <typename T>
static void dfree(wfx_ctx *c, T *&p) {
  if (!p) return;
  auto it = std::find(c->allocations.begin(), c->allocations.end(), (void *)p);
  if (it != c->allocations.end()) {
    (void)hipDeviceSynchronize();  // no queued work may still use it
    (void)hipFree(p);
    c->allocations.erase(it);
  }
  p = nullptr;
}

template <typename T>
static hipError_t dalloc(wfx_ctx *c, T **p, size_t count) {
  void *q = nullptr;
  size_t bytes = (count > 0 ? count : 1) * sizeof(T);
  hipError_t e = hipMalloc(&q, bytes);
  if (e != hipSuccess) return e;
  c->allocations.push_back(q);
  e = hipMemset(q, 0, bytes);
  *p = (T *)q;
  return e;
}

static void drop_graph(wfx_ctx *c) {
  if (c->graph_exec) (void)hipGraphExecDestroy(c->graph_exec);
  if (c->graph) (void)hipGraphDestroy(c->graph);
  c->graph_exec = nullptr;
  c->graph = nullptr;
  c->graph_steps = 0;
}

static void free_all(wfx_ctx *c) {
  drop_graph(c);
  adjoint_free(c);
  frames_free(c);  // pending frames are written first
  for (void *p : c->allocations) (void)hipFree(p);
  c->allocations.clear();
  c->nd = DevNode();
  c->ms = DevMesh();
  c->e64 = DevElem<double>();
  c->e32 = DevElem<float>();
  for (int d = 0; d < 3; ++d) { c->bc_nod[d] = nullptr; c->bc_val[d] = nullptr; c->bc_count[d] = 0; }
  c->react = nullptr;
  c->bc_slot = c->bc_mask = nullptr;
  c->bc_tab = nullptr;
  c->v32 = c->pa32 = c->ud32 = c->fi32 = nullptr;
  c->x_start = nullptr;
  c->x_lo = nullptr;
  c->temp_e = c->temp_buf = nullptr;
  c->contact = wfx::Contact();
  c->has_fe = 0;
  c->fused_stale = false;
  c->scalars = nullptr;
  c->status = nullptr;
  c->cfl_part = nullptr;
  c->cfl_cap_min = c->cfl_nvmax = 0;
  c->hist = nullptr;
  c->hist_cap = 0;
  c->dt_hist = nullptr;
  c->dt_cap = 0;
  c->dt_hist_h.clear();
}

template <typename T>
static hipError_t alloc_elem(wfx_ctx *c, DevElem<T> &el) {
  const size_t E = c->E, n = c->n;
  hipError_t e = hipSuccess;
#define A(ptr, cnt) \
  if (e == hipSuccess) e = dalloc(c, &el.ptr, (cnt));
  A(detJ, E) A(dH, 3 * n * E) A(vol, E) A(vol0, E) A(rho, E) A(rho0, E) A(p, E) A(pl_strain, E)
  A(sigma_y, E) A(sy0, E) A(str_rate, 6 * E) A(rot_rate, 3 * E) A(sigma, 6 * E) A(tau, 6 * E)
  // f_elem is padded to whole 64-element tiles (the fused path may tile it);
  // FP32-mixed fused runs keep 4 floats per element-node there (float4)
  A(eps, 6 * E) A(strain_pl_incr, 6 * E) A(f_elem, (std::is_same<T, float>::value ? 4 : 3) * n * ((E + 63) / 64 * 64))
  A(f_hg, 3 * n * E) A(avq, E) A(part, 4 * (c->ms.nslot > 0 ? c->ms.nslot : 1))
#undef A
  return e;
}

// ---- static scratch of the allocation-free entry point ----------------------
// wfx_energies_noalloc keeps its reduction partials (two doubles per
// 256-thread block over max(N, E)) in this file-scope array instead of a
// buffer it allocates.  Each GPU holds its own copy until the process ends.
// On the host the array's name is only a stand-in, never a device address:
// host code uses what hipGetSymbolAddress returns on each device, kept in
// g_scratch_dev and copied into every context set up on that device.
constexpr long kScratchDoubles = 1L << 17;  // 1 MiB: max(N, E) up to 16,776,960
static __device__ __align__(256) double scratch[kScratchDoubles];
constexpr int kScratchMaxDevices = 64;
static std::mutex g_scratch_mu;
static double *g_scratch_dev[kScratchMaxDevices] = {};

// the current device's address of `scratch`, looked up once per device
static int scratch_address(double **out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) {
    set_error(std::string("hipGetDevice: ") + hipGetErrorString(e));
    return WFX_ERR_HIP;
  }
  if (dev < 0 || dev >= kScratchMaxDevices) {
    set_error("device index " + std::to_string(dev) + " beyond the static scratch table");
    return WFX_ERR_HIP;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      set_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
      return WFX_ERR_HIP;
    }
    g_scratch_dev[dev] = static_cast<double *>(p);
  }
  *out = g_scratch_dev[dev];
  return WFX_OK;
}

// Allocates every device buffer for the mesh held on the host and uploads
// coordinates, connectivity and CSR.
static int setup_device(wfx_ctx *c, const ClusterPlan &plan) {
  free_all(c);
  {
    int rc = scratch_address(&c->scratch);  // the device this context's buffers live on
    if (rc) return rc;
  }
  const size_t N = c->N, E = c->E, n = c->n;
  hipError_t e = hipSuccess;
#define A(ptr, cnt) \
  if (e == hipSuccess) e = dalloc(c, &ptr, (cnt));
  A(c->nd.x, 3 * N) A(c->nd.v, 3 * N) A(c->nd.a, 3 * N) A(c->nd.u, 3 * N) A(c->nd.u_dt, 3 * N)
  A(c->nd.prev_a, 3 * N) A(c->nd.fi, 3 * N) A(c->nd.fe, 3 * N) A(c->nd.mdiag, N) A(c->nd.voln, N)
  A(c->nd.voln0_sum, N) A(c->nd.anp_ratio, N) A(c->nd.T, N)
  A(c->ms.elnod, n * E) A(c->ms.nodel, n * E) A(c->ms.nodel_loc, n * E) A(c->ms.nodel_offset, N)
  A(c->ms.nodel_count, N) A(c->ms.nodel_pk, n * E) A(c->ms.nodel_pad, 8 * N) A(c->bc_slot, N)
  A(c->scalars, 64) A(c->status, 16)
  c->ms.ncl = plan.clusters();
  c->ms.nslot = plan.slots();
  // more than 12 elements per node on average (tetrahedra: about 23): the
  // fused nodal volume sums take several lanes per node
  c->f2_lanes = (long)n * E > 12L * (long)N ? 1 : 0;
  if (plan.max_slots > slot_cap((int)n)) {  // the element pass stages at most slot_cap slots
    set_error("element cluster with " + std::to_string(plan.max_slots) + " nodes");
    free_all(c);
    return WFX_ERR_STATE;
  }
  A(c->ms.cl_eoff, c->ms.ncl + 1) A(c->ms.cl_sbase, c->ms.ncl + 1) A(c->ms.cl_loff, c->ms.nslot + c->ms.ncl)
  A(c->ms.cl_ent, n * E) A(c->ms.nsl_off, N + 1)
  A(c->ms.cl_snode, c->ms.nslot) A(c->ms.cl_lslot, n * E)
  if (c->precision == WFX_FP64) {
    A(c->ms.nsl, c->ms.nslot)
  } else {
    A(c->ms.spos, c->ms.nslot)
  }
  // partial minima of the volume pass (one per element cluster) and of the
  // per-step kernels at block size >= 64
  c->cfl_cap_min = (int)std::max<size_t>((E + 63) / 64 + 1, (size_t)c->ms.ncl + 1);
  A(c->cfl_part, c->cfl_cap_min + (N + 63) / 64 + 1)
  if (e == hipSuccess) e = (c->precision == WFX_FP64) ? alloc_elem(c, c->e64) : alloc_elem(c, c->e32);
  A(c->x_start, 3 * N)
  if (c->precision == WFX_FP32_MIXED) {
    A(c->v32, 3 * N) A(c->pa32, 3 * N) A(c->ud32, 3 * N) A(c->fi32, 3 * N)
  } else {
    A(c->x_lo, 3 * N)
  }
#undef A
  if (e != hipSuccess) {
    set_error(std::string("device allocation: ") + hipGetErrorString(e));
    free_all(c);
    return WFX_ERR_HIP;
  }
  // coordinates [N][3] -> [3][N]; connectivity [E][n] -> [n][E] in the
  // stored (cluster) order; the CSR keeps the reference's order of entries
  // (ascending original element) and names the stored elements
  std::vector<double> xs(3 * N);
  for (size_t i = 0; i < N; ++i)
    for (int d = 0; d < 3; ++d) xs[d * N + i] = c->x_h[3 * i + d];
  std::vector<int> es(n * E);
  for (size_t k = 0; k < E; ++k)
    for (size_t a = 0; a < n; ++a) es[a * E + k] = c->elnod_h[(size_t)c->el_s2o[k] * n + a];
  std::vector<int> nodel_s(n * E);
  for (size_t k = 0; k < n * E; ++k) nodel_s[k] = c->el_o2s[c->nodel_h[k]];
  WFX_CHECK_HIP(hipMemcpy(c->nd.x, xs.data(), 3 * N * sizeof(double), hipMemcpyHostToDevice));
  WFX_CHECK_HIP(hipMemcpy(c->ms.elnod, es.data(), n * E * sizeof(int), hipMemcpyHostToDevice));
  WFX_CHECK_HIP(hipMemcpy(c->ms.nodel, nodel_s.data(), n * E * sizeof(int), hipMemcpyHostToDevice));
  WFX_CHECK_HIP(hipMemcpy(c->ms.nodel_loc, c->nodel_loc_h.data(), n * E * sizeof(int), hipMemcpyHostToDevice));
  WFX_CHECK_HIP(hipMemcpy(c->ms.nodel_offset, c->nodel_offset_h.data(), N * sizeof(int), hipMemcpyHostToDevice));
  WFX_CHECK_HIP(hipMemcpy(c->ms.nodel_count, c->nodel_count_h.data(), N * sizeof(int), hipMemcpyHostToDevice));
  WFX_CHECK_HIP(hipMemcpy(c->ms.cl_eoff, plan.eoff.data(), plan.eoff.size() * sizeof(int), hipMemcpyHostToDevice));
  WFX_CHECK_HIP(hipMemcpy(c->ms.cl_sbase, plan.sbase.data(), plan.sbase.size() * sizeof(int), hipMemcpyHostToDevice));
  WFX_CHECK_HIP(hipMemcpy(c->ms.cl_loff, plan.loff.data(), plan.loff.size() * sizeof(uint16_t), hipMemcpyHostToDevice));
  WFX_CHECK_HIP(hipMemcpy(c->ms.cl_ent, plan.ent.data(), plan.ent.size() * sizeof(uint16_t), hipMemcpyHostToDevice));
  WFX_CHECK_HIP(hipMemcpy(c->ms.nsl_off, plan.nsl_off.data(), plan.nsl_off.size() * sizeof(int), hipMemcpyHostToDevice));
  if (c->ms.nsl)
    WFX_CHECK_HIP(hipMemcpy(c->ms.nsl, plan.nsl.data(), plan.nsl.size() * sizeof(int), hipMemcpyHostToDevice));
  if (c->ms.spos)
    WFX_CHECK_HIP(hipMemcpy(c->ms.spos, plan.spos.data(), plan.spos.size() * sizeof(int), hipMemcpyHostToDevice));
  WFX_CHECK_HIP(hipMemcpy(c->ms.cl_snode, plan.snode.data(), plan.snode.size() * sizeof(int), hipMemcpyHostToDevice));
  WFX_CHECK_HIP(hipMemcpy(c->ms.cl_lslot, plan.lslot.data(), plan.lslot.size() * sizeof(uint16_t),
                          hipMemcpyHostToDevice));
  std::vector<int> pk(n * E);
  for (size_t k = 0; k < n * E; ++k) pk[k] = (nodel_s[k] << 3) | c->nodel_loc_h[k];
  WFX_CHECK_HIP(hipMemcpy(c->ms.nodel_pk, pk.data(), n * E * sizeof(int), hipMemcpyHostToDevice));
  // entry j of node i at j*N + i: the node passes load their first eight
  // entries with unit stride across a wavefront and need no offset for them
  std::vector<int> pad(8 * N, -1);
  for (size_t i = 0; i < N; ++i)
    for (int j = 0; j < std::min(8, c->nodel_count_h[i]); ++j) pad[j * N + i] = pk[c->nodel_offset_h[i] + j];
  WFX_CHECK_HIP(hipMemcpy(c->ms.nodel_pad, pad.data(), 8 * N * sizeof(int), hipMemcpyHostToDevice));
  WFX_CHECK_HIP(hipMemset(c->bc_slot, 0xff, N * sizeof(int)));  // -1: no boundary condition
  WFX_CHECK_HIP(hipMemcpy(c->scalars + kCflDt, &c->prm.dt, sizeof(double), hipMemcpyHostToDevice));
  {
    std::vector<double> t0(N, c->prm.mat.temp);  // setTemp (Domain_d.h:678-683)
    WFX_CHECK_HIP(hipMemcpy(c->nd.T, t0.data(), N * sizeof(double), hipMemcpyHostToDevice));
  }
  c->has_mesh = true;
  c->initialized = false;
  return WFX_OK;
}

static int finish_mesh(wfx_ctx *c) {
  // CSR positions are int and packed entries hold (element << 3) | local node
  if ((int64_t)c->E * c->n > (int64_t)INT32_MAX || (int64_t)c->E >= (int64_t)1 << 28) {
    set_error("meshes are limited to 2^28 elements and 2^31 element-node entries");
    return WFX_ERR_ARG;
  }
  c->nnz = c->E * c->n;
  if (build_nodel(c->N, c->E, c->n, c->elnod_h.data(), c->nodel_h, c->nodel_loc_h, c->nodel_offset_h,
                  c->nodel_count_h)) {
    set_error("element connectivity references a node outside [0, node_count)");
    return WFX_ERR_ARG;
  }
  c->sy0_h.assign(c->E, c->prm.mat.sy0);
  ClusterPlan plan;
  plan_clusters(c->N, c->E, c->n, c->elnod_h.data(), c->x_h.data(), kCluster, plan);
  c->el_s2o.swap(plan.s2o);
  c->el_o2s.swap(plan.o2s);
  int rc = setup_device(c, plan);
  if (rc) return rc;
  c->prm.N = c->N;
  c->prm.E = c->E;
  c->prm.n = c->n;
  return WFX_OK;
}

// host <-> device element-real conversion for the two precisions
template <typename T>
static int upload_elem_real(T *dst, const double *src, size_t count) {
  if constexpr (std::is_same<T, double>::value) {
    WFX_CHECK_HIP(hipMemcpy(dst, src, count * sizeof(double), hipMemcpyHostToDevice));
  } else {
    std::vector<float> tmp(src, src + count);
    WFX_CHECK_HIP(hipMemcpy(dst, tmp.data(), count * sizeof(float), hipMemcpyHostToDevice));
  }
  return WFX_OK;
}

static int upload_real(wfx_ctx *c, void *dst, const double *src, size_t count, bool elem_type) {
  if (!elem_type || c->precision == WFX_FP64) return upload_elem_real<double>((double *)dst, src, count);
  return upload_elem_real<float>((float *)dst, src, count);
}

static int download_real(wfx_ctx *c, double *dst, const void *src, size_t count, bool elem_type) {
  if (!elem_type || c->precision == WFX_FP64) {
    WFX_CHECK_HIP(hipMemcpy(dst, src, count * sizeof(double), hipMemcpyDeviceToHost));
  } else {
    std::vector<float> tmp(count);
    WFX_CHECK_HIP(hipMemcpy(tmp.data(), src, count * sizeof(float), hipMemcpyDeviceToHost));
    for (size_t i = 0; i < count; ++i) dst[i] = tmp[i];
  }
  return WFX_OK;
}

// Element rows [E][rows] between the original order (host) and the stored
// order (device rows [rows][E]): transposed and permuted in one pass.
static void elem_rows_to_device(const wfx_ctx *c, const double *h, int rows, double *dev_rows) {
  const long E = c->E;
  for (long o = 0; o < E; ++o) {
    const long st = c->el_o2s[o];
    for (int r = 0; r < rows; ++r) dev_rows[(long)r * E + st] = h[o * rows + r];
  }
}

// one value per element, original order -> stored order, uploaded
static int upload_elem_values(wfx_ctx *c, void *dst, const double *h, long count) {
  std::vector<double> t((size_t)count);
  elem_rows_to_device(c, h, 1, t.data());
  return upload_real(c, dst, t.data(), t.size(), true);
}

// Description of one transferable field.  Host layout is [cnt][rows] (the
// reference's), device layout is [rows][cnt] starting at dev.
struct FieldDesc {
  void *dev = nullptr;
  long cnt = 0;    // items (nodes, elements, or entries)
  int rows = 1;    // components per item on the host
  int is_int = 0;  // int32 field
  int elem = 0;    // element real type (T) rather than double
  int kind = 0;    // 0 plain transpose, 1 spin (3 device rows -> 6 host columns)
  int per_elem = 0;  // one item per element: stored order on the device
  int elem_ids = 0;  // values are element indices (stored on the device)
};

static bool find_field(wfx_ctx *c, const char *name, FieldDesc &f) {
  const long N = c->N, E = c->E, n = c->n;
  const bool d64 = (c->precision == WFX_FP64);
  auto E64 = [&](void *p64, void *p32) { return d64 ? p64 : p32; };
  std::string s(name);
  f = FieldDesc();
  auto node3 = [&](double *p) { f.dev = p; f.cnt = N; f.rows = 3; };
  auto node1 = [&](double *p) { f.dev = p; f.cnt = N; f.rows = 1; };
  auto el = [&](void *p, int rows) { f.dev = p; f.cnt = E; f.rows = rows; f.elem = 1; f.per_elem = 1; };
  if (s == "x") node3(c->nd.x);
  else if (s == "v") node3(c->nd.v);
  else if (s == "a") node3(c->nd.a);
  else if (s == "u") node3(c->nd.u);
  else if (s == "u_dt") node3(c->nd.u_dt);
  else if (s == "prev_a") node3(c->nd.prev_a);
  else if (s == "fi") node3(c->nd.fi);
  else if (s == "fe") node3(c->nd.fe);
  else if (s == "mdiag") node1(c->nd.mdiag);
  else if (s == "voln") node1(c->nd.voln);
  else if (s == "elnod") { f.dev = c->ms.elnod; f.cnt = E; f.rows = (int)n; f.is_int = 1; f.per_elem = 1; }
  else if (s == "nodel") { f.dev = c->ms.nodel; f.cnt = E * n; f.is_int = 1; f.elem_ids = 1; }
  else if (s == "nodel_loc") { f.dev = c->ms.nodel_loc; f.cnt = E * n; f.is_int = 1; }
  else if (s == "nodel_offset") { f.dev = c->ms.nodel_offset; f.cnt = N; f.is_int = 1; }
  else if (s == "nodel_count") { f.dev = c->ms.nodel_count; f.cnt = N; f.is_int = 1; }
  else if (s == "detJ") el(E64(c->e64.detJ, c->e32.detJ), 1);
  else if (s == "dHdx" || s == "dHdy" || s == "dHdz") {
    long d = s[3] - 'x';
    void *base = d64 ? (void *)(c->e64.dH + d * n * E) : (void *)(c->e32.dH + d * n * E);
    el(base, (int)n);
  }
  else if (s == "vol") el(E64(c->e64.vol, c->e32.vol), 1);
  else if (s == "vol0") el(E64(c->e64.vol0, c->e32.vol0), 1);
  else if (s == "rho") el(E64(c->e64.rho, c->e32.rho), 1);
  else if (s == "rho0") el(E64(c->e64.rho0, c->e32.rho0), 1);
  else if (s == "p") el(E64(c->e64.p, c->e32.p), 1);
  else if (s == "pl_strain") el(E64(c->e64.pl_strain, c->e32.pl_strain), 1);
  else if (s == "sigma_y") el(E64(c->e64.sigma_y, c->e32.sigma_y), 1);
  else if (s == "sy0") el(E64(c->e64.sy0, c->e32.sy0), 1);
  else if (s == "str_rate") el(E64(c->e64.str_rate, c->e32.str_rate), 6);
  else if (s == "rot_rate") { el(E64(c->e64.rot_rate, c->e32.rot_rate), 6); f.kind = 1; }
  else if (s == "sigma") el(E64(c->e64.sigma, c->e32.sigma), 6);
  else if (s == "tau") el(E64(c->e64.tau, c->e32.tau), 6);
  else if (s == "eps") el(E64(c->e64.eps, c->e32.eps), 6);
  else if (s == "strain_pl_incr") el(E64(c->e64.strain_pl_incr, c->e32.strain_pl_incr), 6);
  else if (s == "f_elem") el(E64(c->e64.f_elem, c->e32.f_elem), (int)(3 * n));
  else if (s == "f_hg") el(E64(c->e64.f_hg, c->e32.f_hg), (int)(3 * n));
  else if (s == "T") node1(c->nd.T);
  else if (c->contact.on && (s == "contforce" || s == "ut_prev"))
    node3(s == "contforce" ? c->contact.contforce : c->contact.ut_prev);
  else if (c->contact.on && s == "node_area") node1(c->contact.node_area);
  else if (c->contact.on && s == "cfn") node1(c->contact.cfn);
  else if (c->contact.on && s == "q_cont_conv") node1(c->contact.q_cont);
  else if (c->contact.on && s == "mesh_in_contact") { f.dev = c->contact.mesh_in_contact; f.cnt = N; f.is_int = 1; }
  else if (c->contact.on && s == "elem_length") { f.dev = c->contact.elem_length; f.cnt = E; f.per_elem = 1; }
  else if (c->contact.on && s == "rigid_x") { f.dev = c->contact.rnode; f.cnt = 3L * c->contact.rn; }
  else if (c->contact.on && s == "rigid_normal") { f.dev = c->contact.rnormal; f.cnt = 3L * c->contact.rt; }
  else if (c->contact.on && s == "rigid_pplane") { f.dev = c->contact.rpplane; f.cnt = c->contact.rt; }
  else if (s == "q_plheat" && E64(c->e64.q_plheat, c->e32.q_plheat)) el(E64(c->e64.q_plheat, c->e32.q_plheat), 1);
  // thermal expansion: the element-node temperature rates at the flat
  // indices below N (the reference's m_dTedt[0..N)), and the fused path's
  // per-element means of the last step
  else if (s == "dTedt_flat" && E64(c->e64.dte_flat, c->e32.dte_flat)) {
    f.dev = E64(c->e64.dte_flat, c->e32.dte_flat);
    f.cnt = N;
    f.elem = 1;
  }
  else if (s == "texp" && E64(c->e64.texp, c->e32.texp)) el(E64(c->e64.texp, c->e32.texp), 1);
  else return false;
  return true;
}

static hipStream_t as_stream(wfx_stream s) { return (hipStream_t)s; }

// The fused nodal pass skips the density loads when every element has the
// same density (the stored value, so the mass sums are unchanged).
// Host scan of the element densities and initial volumes after they are set:
// a uniform density lets the fused pass skip its loads; an initial volume
// that is not finite or below 1e-200 (J_local = vol/vol0 not finite) or a
// density that is not positive and finite sets the guard under which the
// fused pass evaluates every pressure and viscosity term in full.
static int scan_elements(wfx_ctx *c) {
  std::vector<double> r(c->E), v0(c->E);
  const bool f64 = c->precision == WFX_FP64;
  int rc = download_real(c, r.data(), f64 ? (void *)c->e64.rho : (void *)c->e32.rho, r.size(), true);
  if (!rc) rc = download_real(c, v0.data(), f64 ? (void *)c->e64.vol0 : (void *)c->e32.vol0, v0.size(), true);
  if (rc) return rc;
  c->rho_uniform = 0.0;
  if (!r.empty() && r[0] > 0.0 &&
      std::all_of(r.begin(), r.end(), [&](double v) { return v == r[0]; }))
    c->rho_uniform = r[0];
  bool bad = false;
  double rmin = DBL_MAX;
  for (long e = 0; e < c->E; ++e) {
    bad = bad || !(r[e] > 0.0 && r[e] <= DBL_MAX) || !(std::fabs(v0[e]) >= 1e-200 && std::fabs(v0[e]) <= DBL_MAX);
    rmin = std::min(rmin, r[e]);
  }
  c->elem_guard = bad ? 1 : 0;
  c->rho_min = (c->E > 0 && !bad) ? rmin : 0.0;
  return WFX_OK;
}

}  // namespace wfx

using namespace wfx;

#define REQUIRE(cond, code, msg) \
  do {                           \
    if (!(cond)) {               \
      set_error(msg);            \
      return code;               \
    }                            \
  } while (0)
#define REQUIRE_CTX(c) REQUIRE((c) != nullptr, WFX_ERR_ARG, "null context")
#define REQUIRE_MESH(c)  \
  REQUIRE_CTX(c);        \
  REQUIRE((c)->has_mesh, WFX_ERR_STATE, "no mesh: call wfx_add_box_length or wfx_set_mesh first")
#define LAUNCH(expr)                                                        \
  do {                                                                      \
    hipError_t _e = (expr);                                                 \
    if (_e != hipSuccess) {                                                 \
      set_error(std::string(#expr) + ": " + hipGetErrorString(_e));        \
      return WFX_ERR_HIP;                                                   \
    }                                                                       \
  } while (0)
#define BY_PRECISION(c, fn, ...) \
  ((c)->precision == WFX_FP64 ? fn<double>(__VA_ARGS__) : fn<float>(__VA_ARGS__))

extern "C" {

const char *wfx_last_error(void) { return g_err.c_str(); }
int wfx_version(void) { return 1; }

int wfx_create(wfx_ctx **out, int precision) {
  REQUIRE(out != nullptr, WFX_ERR_ARG, "null output pointer");
  REQUIRE(precision == WFX_FP64 || precision == WFX_FP32_MIXED, WFX_ERR_ARG, "unknown precision");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
    set_error("no HIP device available");
    return WFX_ERR_HIP;
  }
  wfx_ctx *c = new wfx_ctx();
  c->precision = precision;
  double rho_b = 0.818200;  // Solver_explicit.C:193-197
  c->prm.alpha = (2.0 * rho_b - 1.0) / (1.0 + rho_b);
  c->prm.beta = (5.0 - 3.0 * rho_b) / ((1.0 + rho_b) * (1.0 + rho_b) * (2.0 - rho_b));
  c->prm.gamma = 1.5 - c->prm.alpha;
  *out = c;
  return WFX_OK;
}

int wfx_destroy(wfx_ctx *c) {
  if (!c) return WFX_OK;
  free_all(c);
  delete c;
  return WFX_OK;
}

int wfx_add_box_length(wfx_ctx *c, const double V[3], const double L[3], double r, int red_int,
                       int tritetra) {
  REQUIRE_CTX(c);
  REQUIRE(V && L && r > 0.0, WFX_ERR_ARG, "bad box arguments");
  REQUIRE(L[2] > 0.0, WFX_ERR_ARG, "2D boxes are not supported (L.z must be > 0)");
  REQUIRE(red_int, WFX_ERR_ARG, "only reduced integration is supported");
  frames_join(c);  // pending frame writers read the host connectivity
  int nodxelem = 0;
  if (box_length_mesh(V, L, r, tritetra != 0, c->x_h, c->elnod_h, nodxelem)) {
    set_error("box mesh: no elements or sizes beyond int32");
    return WFX_ERR_ARG;
  }
  c->n = nodxelem;
  c->N = (int)(c->x_h.size() / 3);
  c->E = (int)(c->elnod_h.size() / nodxelem);
  return finish_mesh(c);
}

int wfx_set_mesh(wfx_ctx *c, int node_count, int elem_count, int nodxelem, const double *x,
                 const int *elnod) {
  REQUIRE_CTX(c);
  REQUIRE(node_count > 0 && elem_count > 0 && x && elnod, WFX_ERR_ARG, "bad mesh arguments");
  REQUIRE(nodxelem == 4 || nodxelem == 8, WFX_ERR_ARG, "nodxelem must be 4 (tet) or 8 (hexa)");
  REQUIRE((long)elem_count * nodxelem < 2147483647L, WFX_ERR_ARG, "mesh too large for int32 CSR");
  frames_join(c);  // pending frame writers read the host connectivity
  c->n = nodxelem;
  c->N = node_count;
  c->E = elem_count;
  c->x_h.assign(x, x + 3L * node_count);
  c->elnod_h.assign(elnod, elnod + (long)elem_count * nodxelem);
  return finish_mesh(c);
}

int wfx_get_sizes(wfx_ctx *c, int *N, int *E, int *n) {
  REQUIRE_CTX(c);
  if (N) *N = c->N;
  if (E) *E = c->E;
  if (n) *n = c->n;
  return WFX_OK;
}

int wfx_set_density(wfx_ctx *c, double rho) {
  REQUIRE_MESH(c);
  std::vector<double> h(c->E, rho);
  void *dst = c->precision == WFX_FP64 ? (void *)c->e64.rho0 : (void *)c->e32.rho0;
  return upload_real(c, dst, h.data(), h.size(), true);
}

// Elastic_ (Material.cuh:22-27) and the fields every law shares.  The
// temperature and the strain-rate cap are domain state (Domain_d::T,
// m_max_edot) and survive a change of material.
static Material elastic_material(const wfx_ctx *c, int model, double E, double nu, double rho0,
                                 double sy0) {
  Material m;
  m.model = model;
  m.E = E;
  m.nu = nu;
  m.K = E / (3.0 * (1.0 - 2.0 * nu));
  m.G = E / (2.0 * (1.0 + nu));
  m.rho0 = rho0;
  m.cs0 = std::sqrt(m.K / rho0);
  m.sy0 = sy0;
  m.temp = c->prm.mat.temp;
  m.max_edot = c->prm.mat.max_edot;
  return m;
}

// AssignMaterial (Domain_d.C:945): every element starts at the law's sy0
static int assign_material(wfx_ctx *c, const Material &m) {
  c->prm.mat = m;
  if (c->has_mesh) {
    c->sy0_h.assign(c->E, m.sy0);
    void *dst = c->precision == WFX_FP64 ? (void *)c->e64.sy0 : (void *)c->e32.sy0;
    return upload_elem_values(c, dst, c->sy0_h.data(), (long)c->sy0_h.size());
  }
  return WFX_OK;
}

#define REQUIRE_ELASTIC(E, nu, rho0) \
  REQUIRE((E) > 0.0 && (nu) > -1.0 && (nu) < 0.5 && (rho0) > 0.0, WFX_ERR_ARG, "bad elastic constants")

int wfx_set_material(wfx_ctx *c, int model, double E, double nu, double rho0, double sy0, double p1,
                     double p2) {
  REQUIRE_CTX(c);
  REQUIRE(model == WFX_BILINEAR || model == WFX_HOLLOMON, WFX_ERR_ARG,
          "unknown material model (Johnson-Cook and GMT have setters of their own)");
  REQUIRE_ELASTIC(E, nu, rho0);
  Material m = elastic_material(c, model, E, nu, rho0, sy0);
  if (model == WFX_BILINEAR) {
    REQUIRE(p1 >= 0.0 && p1 < E, WFX_ERR_ARG, "bilinear tangent modulus must be in [0, E)");
    m.Et = p1;
    m.H = E * p1 / (E - p1);
  } else {
    REQUIRE(p1 > 0.0 && p2 > 0.0, WFX_ERR_ARG, "Hollomon K and m must be positive");
    m.Kh = p1;
    m.mh = p2;
    m.eps0 = sy0 / E;  // Material.cuh:95-98
    m.eps1 = std::pow(sy0 / p1, 1. / p2);
  }
  return assign_material(c, m);
}

int wfx_set_material_jc(wfx_ctx *c, double E, double nu, double rho0, double A, double B, double n,
                        double C, double eps_0, double m_, double T_m, double T_t) {
  REQUIRE_CTX(c);
  REQUIRE_ELASTIC(E, nu, rho0);
  REQUIRE(eps_0 > 0.0 && T_m != T_t, WFX_ERR_ARG,
          "Johnson-Cook needs eps_0 > 0 and T_m != T_t");
  Material m = elastic_material(c, WFX_JOHNSON_COOK, E, nu, rho0, A);
  m.B = B;
  m.n = n;
  m.C = C;
  m.eps_0 = eps_0;
  m.m = m_;
  m.T_m = T_m;
  m.T_t = T_t;
  return assign_material(c, m);
}

int wfx_set_material_gmt(wfx_ctx *c, double E, double nu, double rho0, double sy0, const double k[8],
                         const double r[6]) {
  REQUIRE_CTX(c);
  REQUIRE_ELASTIC(E, nu, rho0);
  REQUIRE(k && r, WFX_ERR_ARG, "GMT needs 8 constants and 6 range bounds");
  Material m = elastic_material(c, WFX_GMT, E, nu, rho0, sy0);
  m.n1 = k[0];
  m.n2 = k[1];
  m.C1 = k[2];
  m.C2 = k[3];
  m.m1 = k[4];
  m.m2 = k[5];
  m.I1 = k[6];
  m.I2 = k[7];
  m.e_min = r[0];
  m.e_max = r[1];
  m.er_min = r[2];
  m.er_max = r[3];
  m.T_min = r[4];
  m.T_max = r[5];
  return assign_material(c, m);
}

int wfx_set_temperature(wfx_ctx *c, double T) {
  REQUIRE_CTX(c);
  REQUIRE(std::isfinite(T), WFX_ERR_ARG, "temperature must be finite");
  c->prm.mat.temp = T;
  if (c->has_mesh) {  // setTemp: every node as well
    std::vector<double> t0(c->N, T);
    WFX_CHECK_HIP(hipMemcpy(c->nd.T, t0.data(), t0.size() * sizeof(double), hipMemcpyHostToDevice));
  }
  return WFX_OK;
}

// ---- rigid-surface contact --------------------------------------------------
// SearchExtNodes (Domain_d.C:110-191) for tetrahedra: every face in element
// and tetra_faces order (Domain_d.h:184-186), deduplicated by node set; the
// faces met once are external and keep that order and their element's node
// order.  External nodes list their external faces in face order.
static void external_faces(const wfx_ctx *c, std::vector<int> &faces, std::vector<int> &ext_nodes,
                           std::vector<int> &afc_off, std::vector<int> &afc_face) {
  static const int tf[4][3] = {{0, 1, 2}, {0, 1, 3}, {1, 2, 3}, {0, 2, 3}};
  const int E = c->E, N = c->N;
  std::vector<std::pair<std::array<int, 3>, int>> key((size_t)4 * E);
  for (int e = 0; e < E; ++e)
    for (int f = 0; f < 4; ++f) {
      std::array<int, 3> k = {c->elnod_h[4 * e + tf[f][0]], c->elnod_h[4 * e + tf[f][1]],
                              c->elnod_h[4 * e + tf[f][2]]};
      std::sort(k.begin(), k.end());
      key[4 * e + f] = {k, 4 * e + f};
    }
  std::sort(key.begin(), key.end());
  std::vector<int> cnt((size_t)4 * E, 0);
  for (size_t a = 0; a < key.size();) {
    size_t b = a;
    while (b < key.size() && key[b].first == key[a].first) ++b;
    for (size_t q = a; q < b; ++q) cnt[key[q].second] = (int)(b - a);
    a = b;
  }
  faces.clear();
  std::vector<std::vector<int>> nf((size_t)N);
  for (int g = 0; g < 4 * E; ++g) {
    if (cnt[g] != 1) continue;
    const int e = g / 4, f = g % 4, id = (int)(faces.size() / 3);
    for (int k = 0; k < 3; ++k) {
      const int nd = c->elnod_h[4 * e + tf[f][k]];
      faces.push_back(nd);
      nf[nd].push_back(id);
    }
  }
  ext_nodes.clear();
  afc_off.assign(1, 0);
  afc_face.clear();
  for (int i = 0; i < N; ++i) {
    if (nf[i].empty()) continue;
    ext_nodes.push_back(i);
    afc_face.insert(afc_face.end(), nf[i].begin(), nf[i].end());
    afc_off.push_back((int)afc_face.size());
  }
}

int wfx_set_rigid_mesh(wfx_ctx *c, int node_count, const double *nodes, const double *node_vel,
                       int tri_count, const int *tris, const int *tri_mesh_id) {
  REQUIRE_MESH(c);
  REQUIRE(c->n == 4, WFX_ERR_ARG,
          "contact needs a tetrahedral body (the external-face search uses tetrahedron faces)");
  REQUIRE(node_count > 0 && tri_count > 0 && nodes && tris, WFX_ERR_ARG, "empty rigid surface");
  for (int k = 0; k < 3 * tri_count; ++k)
    REQUIRE(tris[k] >= 0 && tris[k] < node_count, WFX_ERR_ARG, "rigid triangle node out of range");
  std::vector<int> faces, ext, off, afc;
  external_faces(c, faces, ext, off, afc);
  REQUIRE(!ext.empty(), WFX_ERR_ARG, "the body has no external face");
  auto &ct = c->contact;
  const size_t N = c->N;
  drop_graph(c);
  // a new surface replaces the previous one
  dfree(c, ct.rnode); dfree(c, ct.rvel); dfree(c, ct.rnormal); dfree(c, ct.rpplane); dfree(c, ct.rtri);
  dfree(c, ct.rmesh); dfree(c, ct.ext_nodes); dfree(c, ct.afc_off); dfree(c, ct.afc_face);
  dfree(c, ct.ext_face); dfree(c, ct.node_area); dfree(c, ct.elem_length); dfree(c, ct.contforce);
  dfree(c, ct.ut_prev); dfree(c, ct.mesh_in_contact); dfree(c, ct.cfn); dfree(c, ct.q_cont);
  ct.rn = node_count;
  ct.rt = tri_count;
  ct.next = (int)ext.size();
  WFX_CHECK_HIP(dalloc(c, &ct.rnode, 3 * (size_t)node_count));
  WFX_CHECK_HIP(dalloc(c, &ct.rvel, 3 * (size_t)node_count));
  WFX_CHECK_HIP(dalloc(c, &ct.rnormal, 3 * (size_t)tri_count));
  WFX_CHECK_HIP(dalloc(c, &ct.rpplane, (size_t)tri_count));
  WFX_CHECK_HIP(dalloc(c, &ct.rtri, 3 * (size_t)tri_count));
  WFX_CHECK_HIP(dalloc(c, &ct.rmesh, (size_t)tri_count));
  WFX_CHECK_HIP(dalloc(c, &ct.ext_nodes, ext.size()));
  WFX_CHECK_HIP(dalloc(c, &ct.afc_off, off.size()));
  WFX_CHECK_HIP(dalloc(c, &ct.afc_face, afc.size()));
  WFX_CHECK_HIP(dalloc(c, &ct.ext_face, faces.size()));
  WFX_CHECK_HIP(dalloc(c, &ct.node_area, N));
  WFX_CHECK_HIP(dalloc(c, &ct.elem_length, (size_t)c->E));
  WFX_CHECK_HIP(dalloc(c, &ct.contforce, 3 * N));
  WFX_CHECK_HIP(dalloc(c, &ct.ut_prev, 3 * N));
  WFX_CHECK_HIP(dalloc(c, &ct.mesh_in_contact, N));
  WFX_CHECK_HIP(dalloc(c, &ct.cfn, N));
  WFX_CHECK_HIP(dalloc(c, &ct.q_cont, N));
  std::vector<double> vel(3 * (size_t)node_count, 0.0);
  if (node_vel) std::copy(node_vel, node_vel + vel.size(), vel.begin());
  std::vector<int> mid((size_t)tri_count, 0);
  if (tri_mesh_id) std::copy(tri_mesh_id, tri_mesh_id + tri_count, mid.begin());
  auto up = [](void *dst, const void *src, size_t bytes) { return hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice); };
  WFX_CHECK_HIP(up(ct.rnode, nodes, 3 * sizeof(double) * node_count));
  WFX_CHECK_HIP(up(ct.rvel, vel.data(), vel.size() * sizeof(double)));
  WFX_CHECK_HIP(up(ct.rtri, tris, 3 * sizeof(int) * tri_count));
  WFX_CHECK_HIP(up(ct.rmesh, mid.data(), mid.size() * sizeof(int)));
  WFX_CHECK_HIP(up(ct.ext_nodes, ext.data(), ext.size() * sizeof(int)));
  WFX_CHECK_HIP(up(ct.afc_off, off.data(), off.size() * sizeof(int)));
  WFX_CHECK_HIP(up(ct.afc_face, afc.data(), afc.size() * sizeof(int)));
  WFX_CHECK_HIP(up(ct.ext_face, faces.data(), faces.size() * sizeof(int)));
  WFX_CHECK_HIP(hipMemset(ct.node_area, 0, N * sizeof(double)));
  WFX_CHECK_HIP(hipMemset(ct.elem_length, 0, (size_t)c->E * sizeof(double)));
  WFX_CHECK_HIP(hipMemset(ct.contforce, 0, 3 * N * sizeof(double)));
  WFX_CHECK_HIP(hipMemset(ct.ut_prev, 0, 3 * N * sizeof(double)));
  WFX_CHECK_HIP(hipMemset(ct.cfn, 0, N * sizeof(double)));
  WFX_CHECK_HIP(hipMemset(ct.q_cont, 0, N * sizeof(double)));
  WFX_CHECK_HIP(hipMemset(ct.mesh_in_contact, 0xff, N * sizeof(int)));
  ct.on = 1;
  // normals and plane offsets (a zero-length move), areas of the external faces
  LAUNCH(contact_rigid_update(c, 0.0, nullptr));
  LAUNCH(contact_face_areas(c, nullptr));
  LAUNCH(contact_elem_length(c, nullptr));
  WFX_CHECK_HIP(hipDeviceSynchronize());
  return WFX_OK;
}

int wfx_calc_ext_face_areas(wfx_ctx *c, wfx_stream s) {
  REQUIRE_MESH(c);
  REQUIRE(c->contact.on, WFX_ERR_STATE, "no rigid surface: call wfx_set_rigid_mesh");
  LAUNCH(contact_face_areas(c, as_stream(s)));
  return WFX_OK;
}

int wfx_calc_contact_forces(wfx_ctx *c, wfx_stream s) {
  REQUIRE_MESH(c);
  REQUIRE(c->contact.on, WFX_ERR_STATE, "no rigid surface: call wfx_set_rigid_mesh");
  REQUIRE(c->initialized, WFX_ERR_STATE, "call wfx_init first");
  LAUNCH(contact_forces(c, as_stream(s)));
  return WFX_OK;
}

int wfx_rigid_update(wfx_ctx *c, wfx_stream s) {
  REQUIRE_MESH(c);
  REQUIRE(c->contact.on, WFX_ERR_STATE, "no rigid surface: call wfx_set_rigid_mesh");
  LAUNCH(contact_rigid_update(c, c->prm.dt, as_stream(s)));
  return WFX_OK;
}

int wfx_set_contact_params(wfx_ctx *c, double penalty_factor, double mu_static, double mu_dynamic) {
  REQUIRE_CTX(c);
  REQUIRE(penalty_factor >= 0.0 && mu_static >= 0.0 && mu_dynamic >= 0.0, WFX_ERR_ARG,
          "penalty factor and friction coefficients must be non-negative");
  c->contact.pf = penalty_factor;
  c->contact.mu_sta = mu_static;
  c->contact.mu_dyn = mu_dynamic;
  return WFX_OK;
}

int wfx_set_contact_heat(wfx_ctx *c, double heat_cond, double die_temp) {
  REQUIRE_CTX(c);
  REQUIRE(heat_cond >= 0.0 && std::isfinite(heat_cond) && std::isfinite(die_temp), WFX_ERR_ARG,
          "the conduction coefficient must be non-negative and finite, the die temperature finite");
  c->contact.heat_cond = heat_cond;
  c->contact.T_const = die_temp;
  return WFX_OK;
}

int wfx_set_rigid_velocity(wfx_ctx *c, const double *node_vel) {
  REQUIRE_MESH(c);
  REQUIRE(c->contact.on, WFX_ERR_STATE, "no rigid surface: call wfx_set_rigid_mesh");
  REQUIRE(node_vel, WFX_ERR_ARG, "null velocities");
  WFX_CHECK_HIP(hipMemcpy(c->contact.rvel, node_vel, 3 * sizeof(double) * c->contact.rn, hipMemcpyHostToDevice));
  return WFX_OK;
}

int wfx_contact_force(wfx_ctx *c, int mesh_id, double *out) {
  REQUIRE_MESH(c);
  REQUIRE(c->contact.on, WFX_ERR_STATE, "no rigid surface: call wfx_set_rigid_mesh");
  REQUIRE(out, WFX_ERR_ARG, "null output");
  std::vector<int> mic(c->N);
  std::vector<double> cfn(c->N);
  WFX_CHECK_HIP(hipDeviceSynchronize());
  WFX_CHECK_HIP(hipMemcpy(mic.data(), c->contact.mesh_in_contact, mic.size() * sizeof(int), hipMemcpyDeviceToHost));
  WFX_CHECK_HIP(hipMemcpy(cfn.data(), c->contact.cfn, cfn.size() * sizeof(double), hipMemcpyDeviceToHost));
  double r = 0.0;  // node order, as the reference's serial sum (Contact.C:318-321)
  for (int i = 0; i < c->N; ++i)
    if (mic[i] == mesh_id) r += cfn[i];
  *out = r;
  return WFX_OK;
}

int wfx_set_thermal(wfx_ctx *c, int on, double k_T, double cp_T, double plheat_fraction) {
  REQUIRE_MESH(c);
  if (on) {
    REQUIRE(k_T >= 0.0 && cp_T > 0.0 && plheat_fraction >= 0.0 && std::isfinite(k_T) && std::isfinite(cp_T),
            WFX_ERR_ARG, "thermal coupling needs k_T >= 0, cp_T > 0 and a non-negative heat fraction");
    // element-node rates in the element-force layout: whole 64-element tiles
    const size_t E = c->E, n = c->n, Ep = (E + 63) / 64 * 64;
    if (c->precision == WFX_FP64) {
      if (!c->e64.q_plheat) WFX_CHECK_HIP(dalloc(c, &c->e64.q_plheat, E));
      if (!c->e64.dTedt) WFX_CHECK_HIP(dalloc(c, &c->e64.dTedt, n * Ep));
    } else {
      if (!c->e32.q_plheat) WFX_CHECK_HIP(dalloc(c, &c->e32.q_plheat, E));
      if (!c->e32.dTedt) WFX_CHECK_HIP(dalloc(c, &c->e32.dTedt, n * Ep));
    }
    void *q = c->precision == WFX_FP64 ? (void *)c->e64.q_plheat : (void *)c->e32.q_plheat;
    WFX_CHECK_HIP(hipMemset(q, 0, E * (c->precision == WFX_FP64 ? sizeof(double) : sizeof(float))));
  }
  c->thermal.on = on ? 1 : 0;
  c->thermal.k_T = k_T;
  c->thermal.cp_T = cp_T;
  c->thermal.plheat_frac = plheat_fraction;
  return WFX_OK;
}

// Material_::exp_T ("thermalExp", main.C:577): with thermal coupling on,
// every step removes exp_T times the element's temperature rate from the
// diagonal strain rate (calcThermalExpansion, Thermal.C:164-178).  The rates
// it reads are the last step's element-node rates at the flat indices
// [e*n + a] equal to the element's node ids, as the reference reads them:
// the library keeps those below N (flat_base maps each stored element to
// e*n of its original index) and starts them at zero.
int wfx_set_thermal_expansion(wfx_ctx *c, double exp_T) {
  REQUIRE_MESH(c);
  REQUIRE(std::isfinite(exp_T), WFX_ERR_ARG, "thermal expansion coefficient must be finite");
  if (exp_T != 0.0) {
    const size_t E = c->E, N = c->N;
    if (!c->ms.flat_base) {
      WFX_CHECK_HIP(dalloc(c, &c->ms.flat_base, E));
      std::vector<int> fb(E);
      for (size_t st = 0; st < E; ++st) fb[st] = c->el_s2o[st] * c->n;
      WFX_CHECK_HIP(hipMemcpy(c->ms.flat_base, fb.data(), E * sizeof(int), hipMemcpyHostToDevice));
    }
    if (c->precision == WFX_FP64) {
      if (!c->e64.dte_flat) WFX_CHECK_HIP(dalloc(c, &c->e64.dte_flat, N));
      if (!c->e64.texp) WFX_CHECK_HIP(dalloc(c, &c->e64.texp, E));
    } else {
      if (!c->e32.dte_flat) WFX_CHECK_HIP(dalloc(c, &c->e32.dte_flat, N));
      if (!c->e32.texp) WFX_CHECK_HIP(dalloc(c, &c->e32.texp, E));
    }
  }
  c->thermal.exp_T = exp_T;
  drop_graph(c);
  return WFX_OK;
}

int wfx_calc_thermal_expansion(wfx_ctx *c, wfx_stream s) {
  REQUIRE_MESH(c);
  REQUIRE(!c->fused_stale, WFX_ERR_STATE, "calcThermalExpansion needs the strain rates of a per-step API step");
  LAUNCH(BY_PRECISION(c, launch_thermal_expansion, c, as_stream(s)));
  return WFX_OK;
}

int wfx_thermal_calcs(wfx_ctx *c, wfx_stream s) {
  REQUIRE_MESH(c);
  REQUIRE(c->thermal.on, WFX_ERR_STATE, "thermal coupling is off: call wfx_set_thermal");
  REQUIRE(!c->fused_stale, WFX_ERR_STATE, "ThermalCalcs needs the Jacobians of a per-step API step");
  LAUNCH(BY_PRECISION(c, launch_thermal, c, as_stream(s)));
  return WFX_OK;
}

int wfx_set_elem_temperature(wfx_ctx *c, const double *T, int count) {
  REQUIRE_MESH(c);
  if (!T) {
    c->temp_e = nullptr;  // the buffer stays allocated until the mesh changes
    return WFX_OK;
  }
  REQUIRE(count == c->E, WFX_ERR_ARG, "temperature needs one value per element");
  if (!c->temp_buf) WFX_CHECK_HIP(dalloc(c, &c->temp_buf, c->E));
  {
    std::vector<double> t((size_t)count);
    elem_rows_to_device(c, T, 1, t.data());
    WFX_CHECK_HIP(hipMemcpy(c->temp_buf, t.data(), sizeof(double) * (size_t)count, hipMemcpyHostToDevice));
  }
  c->temp_e = c->temp_buf;
  return WFX_OK;
}

int wfx_set_max_strain_rate(wfx_ctx *c, double edot) {
  REQUIRE_CTX(c);
  REQUIRE(edot >= 0.0, WFX_ERR_ARG, "strain-rate cap must be non-negative");
  c->prm.mat.max_edot = edot;
  return WFX_OK;
}

int wfx_set_elem_sy0(wfx_ctx *c, const double *sy0, int count) {
  REQUIRE_MESH(c);
  REQUIRE(sy0 && count == c->E, WFX_ERR_ARG, "sy0 needs one value per element");
  c->sy0_h.assign(sy0, sy0 + count);
  void *dst = c->precision == WFX_FP64 ? (void *)c->e64.sy0 : (void *)c->e32.sy0;
  return upload_elem_values(c, dst, sy0, count);
}

int wfx_add_bc_vel_node(wfx_ctx *c, int node, int dim, double val) {
  REQUIRE_MESH(c);
  REQUIRE(dim >= 0 && dim < 3, WFX_ERR_ARG, "dim must be 0, 1 or 2");
  REQUIRE(node >= 0 && node < c->N, WFX_ERR_ARG, "node out of range");
  c->bc_nod_h[dim].push_back(node);
  c->bc_val_h[dim].push_back(val);
  return WFX_OK;
}

int wfx_allocate_bcs(wfx_ctx *c) {
  REQUIRE_MESH(c);
  for (int d = 0; d < 3; ++d) {
    // the reference applies its list serially, so for a node listed twice the
    // last value wins; keep exactly that entry (first-occurrence order kept)
    std::vector<int> last(c->N, -1);
    const auto &nod = c->bc_nod_h[d];
    for (size_t k = 0; k < nod.size(); ++k) last[nod[k]] = (int)k;
    std::vector<int> dn;
    std::vector<double> dv;
    for (size_t k = 0; k < nod.size(); ++k)
      if (last[nod[k]] >= 0) {
        dn.push_back(nod[k]);
        dv.push_back(c->bc_val_h[d][last[nod[k]]]);
        last[nod[k]] = -1;
      }
    int cnt = (int)dn.size();
    if (cnt > 0) {
      int *pn = nullptr;
      double *pv = nullptr;
      WFX_CHECK_HIP(dalloc(c, &pn, cnt));
      WFX_CHECK_HIP(dalloc(c, &pv, cnt));
      WFX_CHECK_HIP(hipMemcpy(pn, dn.data(), cnt * sizeof(int), hipMemcpyHostToDevice));
      WFX_CHECK_HIP(hipMemcpy(pv, dv.data(), cnt * sizeof(double), hipMemcpyHostToDevice));
      c->bc_nod[d] = pn;
      c->bc_val[d] = pv;
    }
    c->bc_count[d] = cnt;
  }
  // per-node view of the same data for the fused node pass
  std::vector<int> slot(c->N, -1), mask;
  std::vector<double> tab;
  for (int d = 0; d < 3; ++d) {
    const auto &nod = c->bc_nod_h[d];
    for (size_t k = 0; k < nod.size(); ++k) {
      int g = nod[k];
      if (slot[g] < 0) {
        slot[g] = (int)mask.size();
        mask.push_back(0);
        tab.insert(tab.end(), {0.0, 0.0, 0.0});
      }
      mask[slot[g]] |= 1 << d;
      tab[3 * slot[g] + d] = c->bc_val_h[d][k];  // last entry wins
    }
  }
  if (!mask.empty()) {
    int *pm = nullptr;
    double *pt = nullptr;
    WFX_CHECK_HIP(dalloc(c, &pm, mask.size()));
    WFX_CHECK_HIP(dalloc(c, &pt, tab.size()));
    WFX_CHECK_HIP(hipMemcpy(pm, mask.data(), mask.size() * sizeof(int), hipMemcpyHostToDevice));
    WFX_CHECK_HIP(hipMemcpy(pt, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice));
    c->bc_mask = pm;
    c->bc_tab = pt;
  }
  WFX_CHECK_HIP(hipMemcpy(c->bc_slot, slot.data(), c->N * sizeof(int), hipMemcpyHostToDevice));
  return WFX_OK;
}

int wfx_set_dt(wfx_ctx *c, double dt) {
  REQUIRE_CTX(c);
  REQUIRE(dt > 0.0 && std::isfinite(dt), WFX_ERR_ARG, "dt must be positive");
  c->prm.dt = dt;
  // the device copy the first prediction of a per-step-CFL run reads
  if (c->scalars) WFX_CHECK_HIP(hipMemcpy(c->scalars + kCflDt, &dt, sizeof(double), hipMemcpyHostToDevice));
  return WFX_OK;
}

int wfx_set_cfl(wfx_ctx *c, double cfl_factor) {
  REQUIRE_CTX(c);
  REQUIRE(cfl_factor >= 0.0 && std::isfinite(cfl_factor), WFX_ERR_ARG, "CFL factor must be >= 0");
  c->prm.cfl = cfl_factor;
  return WFX_OK;
}

int wfx_get_dt(wfx_ctx *c, double *dt) {
  REQUIRE_CTX(c);
  REQUIRE(dt, WFX_ERR_ARG, "null output");
  *dt = c->prm.dt;
  return WFX_OK;
}

int wfx_dt_history(wfx_ctx *c, double *out, int count) {
  REQUIRE_CTX(c);
  REQUIRE(out && count >= 0 && (size_t)count <= c->dt_hist_h.size(), WFX_ERR_ARG,
          "count exceeds the steps of the last run with a per-step CFL time step");
  std::copy(c->dt_hist_h.begin(), c->dt_hist_h.begin() + count, out);
  return WFX_OK;
}

int wfx_set_ch(wfx_ctx *c, double alpha, double beta, double gamma) {
  REQUIRE_CTX(c);
  REQUIRE(alpha != 1.0, WFX_ERR_ARG, "alpha must differ from 1");
  c->prm.alpha = alpha;
  c->prm.beta = beta;
  c->prm.gamma = gamma;
  return WFX_OK;
}

static double *param_slot(wfx_ctx *c, const char *name) {
  Stab &s = c->prm.stab;
  struct { const char *nm; double *p; } t[] = {
      {"alpha_free", &s.alpha_free}, {"log_factor", &s.log_factor}, {"J_min", &s.J_min},
      {"hg_coeff_free", &s.hg_coeff_free}, {"alpha_contact", &s.alpha_contact},
      {"hg_coeff_contact", &s.hg_coeff_contact}, {"pspg_scale", &s.pspg_scale},
      {"p_pspg_bulkfac", &s.p_pspg_bulkfac}, {"av_coeff_div", &s.av_coeff_div},
      {"av_coeff_bulk", &s.av_coeff_bulk}, {"artifvisc0", &s.artifvisc0},
      {"artifvisc1", &s.artifvisc1}, {"hg_coeff", &s.hg_coeff}, {"dt", &c->prm.dt},
      {"alpha", &c->prm.alpha}, {"beta", &c->prm.beta}, {"gamma", &c->prm.gamma}};
  for (auto &e : t)
    if (!std::strcmp(e.nm, name)) return e.p;
  return nullptr;
}

int wfx_set_param(wfx_ctx *c, const char *name, double value) {
  REQUIRE_CTX(c);
  REQUIRE(name, WFX_ERR_ARG, "null name");
  if (!std::strcmp(name, "check_status")) {
    REQUIRE(value == 0.0 || value == 1.0, WFX_ERR_ARG, "check_status is 0 or 1");
    c->check_status = (int)value;
    return WFX_OK;
  }
  double *p = param_slot(c, name);
  REQUIRE(p, WFX_ERR_ARG, std::string("unknown parameter ") + name);
  *p = value;
  return WFX_OK;
}

int wfx_get_param(wfx_ctx *c, const char *name, double *value) {
  REQUIRE_CTX(c);
  REQUIRE(name && value, WFX_ERR_ARG, "null argument");
  std::string s(name);
  const Material &m = c->prm.mat;
  if (s == "K") { *value = m.K; return WFX_OK; }
  if (s == "G") { *value = m.G; return WFX_OK; }
  if (s == "cs0") { *value = m.cs0; return WFX_OK; }
  if (s == "H") { *value = m.H; return WFX_OK; }
  if (s == "check_status") { *value = c->check_status; return WFX_OK; }
  // read-only: the element clusters of the fused element pass
  if (s == "cluster_count") { *value = c->ms.ncl; return WFX_OK; }
  if (s == "slot_count") { *value = (double)c->ms.nslot; return WFX_OK; }
  double *p = param_slot(c, name);
  REQUIRE(p, WFX_ERR_ARG, std::string("unknown parameter ") + name);
  *value = *p;
  return WFX_OK;
}

int wfx_set_reaction_groups(wfx_ctx *c, const int *nodes, int count, const int *group_offsets,
                            int ngroups) {
  REQUIRE_MESH(c);
  REQUIRE(count >= 0 && (count == 0 || nodes), WFX_ERR_ARG, "bad node list");
  REQUIRE(ngroups >= 1 && group_offsets, WFX_ERR_ARG, "need at least one group");
  REQUIRE(group_offsets[0] == 0 && group_offsets[ngroups] == count, WFX_ERR_ARG,
          "group offsets must start at 0 and end at count");
  for (int g = 0; g < ngroups; ++g)
    REQUIRE(group_offsets[g + 1] >= group_offsets[g], WFX_ERR_ARG, "group offsets must not decrease");
  for (int k = 0; k < count; ++k) REQUIRE(nodes[k] >= 0 && nodes[k] < c->N, WFX_ERR_ARG, "reaction node out of range");
  c->react_h.assign(nodes, nodes + count);
  c->react_off_h.assign(group_offsets, group_offsets + ngroups + 1);
  int *p = nullptr, *q = nullptr;
  WFX_CHECK_HIP(dalloc(c, &p, count));
  WFX_CHECK_HIP(dalloc(c, &q, ngroups + 1));
  if (count) WFX_CHECK_HIP(hipMemcpy(p, nodes, count * sizeof(int), hipMemcpyHostToDevice));
  WFX_CHECK_HIP(hipMemcpy(q, group_offsets, (ngroups + 1) * sizeof(int), hipMemcpyHostToDevice));
  c->react = p;
  c->react_off = q;
  c->ngroups = ngroups;
  c->hist_cap = 0;  // history layout changed
  return WFX_OK;
}

int wfx_set_reaction_nodes(wfx_ctx *c, const int *nodes, int count) {
  const int off[2] = {0, count};
  return wfx_set_reaction_groups(c, nodes, count, off, 1);
}

int wfx_reaction_group_count(wfx_ctx *c, int *ngroups) {
  REQUIRE_CTX(c);
  REQUIRE(ngroups, WFX_ERR_ARG, "null output");
  *ngroups = c->ngroups;
  return WFX_OK;
}

int wfx_calc_min_edge_length(wfx_ctx *c, double *out, wfx_stream s) {
  REQUIRE_MESH(c);
  REQUIRE(out, WFX_ERR_ARG, "null output");
  double *partial = nullptr;
  WFX_CHECK_HIP(hipMalloc(&partial, sizeof(double) * (min_edge_partial_count(c) + 1)));
  hipError_t e = launch_min_edge(c, partial, c->scalars + 8, as_stream(s));
  if (e == hipSuccess) e = hipMemcpyAsync(out, c->scalars + 8, sizeof(double), hipMemcpyDeviceToHost, as_stream(s));
  if (e == hipSuccess) e = hipStreamSynchronize(as_stream(s));
  (void)hipFree(partial);
  WFX_CHECK_HIP(e);
  return WFX_OK;
}

// ---- init block -----------------------------------------------------------
int wfx_init(wfx_ctx *c, wfx_stream s) {
  REQUIRE_MESH(c);
  REQUIRE(c->prm.mat.E > 0.0, WFX_ERR_STATE, "no material: call wfx_set_material");
  hipStream_t st = as_stream(s);
  const size_t N = c->N, E = c->E;
  // InitValues (Domain_d.C:393-416) plus zero state; every buffer the
  // reference leaves uninitialised starts at zero here.
  for (double *p : {c->nd.v, c->nd.a, c->nd.u, c->nd.u_dt, c->nd.prev_a, c->nd.fi})
    WFX_CHECK_HIP(hipMemsetAsync(p, 0, 3 * N * sizeof(double), st));
  size_t es = c->precision == WFX_FP64 ? sizeof(double) : sizeof(float);
  auto zero_e = [&](void *p, size_t rows) { return hipMemsetAsync(p, 0, rows * E * es, st); };
  if (c->precision == WFX_FP64) {
    auto &el = c->e64;
    for (auto pr : {std::make_pair((void *)el.p, 1), {(void *)el.pl_strain, 1}, {(void *)el.tau, 6},
                    {(void *)el.eps, 6}, {(void *)el.sigma, 6}, {(void *)el.strain_pl_incr, 6},
                    {(void *)el.f_hg, 3 * c->n}})
      WFX_CHECK_HIP(zero_e(pr.first, pr.second));
    WFX_CHECK_HIP(hipMemcpyAsync(el.sigma_y, el.sy0, E * es, hipMemcpyDeviceToDevice, st));
  } else {
    auto &el = c->e32;
    for (auto pr : {std::make_pair((void *)el.p, 1), {(void *)el.pl_strain, 1}, {(void *)el.tau, 6},
                    {(void *)el.eps, 6}, {(void *)el.sigma, 6}, {(void *)el.strain_pl_incr, 6},
                    {(void *)el.f_hg, 3 * c->n}})
      WFX_CHECK_HIP(zero_e(pr.first, pr.second));
    WFX_CHECK_HIP(hipMemcpyAsync(el.sigma_y, el.sy0, E * es, hipMemcpyDeviceToDevice, st));
  }
  for (int d = 0; d < 3; ++d) LAUNCH(launch_impose(c, d, false, st));
  LAUNCH(BY_PRECISION(c, launch_elem_jacobians, c, st));
  LAUNCH(BY_PRECISION(c, launch_elem_vol, c, st));
  void *vol = c->precision == WFX_FP64 ? (void *)c->e64.vol : (void *)c->e32.vol;
  void *vol0 = c->precision == WFX_FP64 ? (void *)c->e64.vol0 : (void *)c->e32.vol0;
  WFX_CHECK_HIP(hipMemcpyAsync(vol0, vol, E * es, hipMemcpyDeviceToDevice, st));
  LAUNCH(BY_PRECISION(c, launch_elem_density, c, st));
  LAUNCH(BY_PRECISION(c, launch_nodal_vol, c, st));
  LAUNCH(BY_PRECISION(c, launch_nodal_mass, c, st));
  WFX_CHECK_HIP(hipMemsetAsync(c->status, 0, 16 * sizeof(unsigned int), st));
  c->status_seen[0] = c->status_seen[1] = 0;
  {  // the temperature rates thermal expansion reads start at zero
    void *dte = c->precision == WFX_FP64 ? (void *)c->e64.dte_flat : (void *)c->e32.dte_flat;
    if (dte) WFX_CHECK_HIP(hipMemsetAsync(dte, 0, N * (c->precision == WFX_FP64 ? sizeof(double) : sizeof(float)), st));
  }
  if (c->contact.on) {
    // contact state starts clear; calcMinEdgeLength before the solve
    // (main.C:862) gives the element heights
    auto &ct = c->contact;
    WFX_CHECK_HIP(hipMemsetAsync(ct.contforce, 0, 3 * N * sizeof(double), st));
    WFX_CHECK_HIP(hipMemsetAsync(ct.ut_prev, 0, 3 * N * sizeof(double), st));
    WFX_CHECK_HIP(hipMemsetAsync(ct.cfn, 0, N * sizeof(double), st));
    WFX_CHECK_HIP(hipMemsetAsync(ct.mesh_in_contact, 0xff, N * sizeof(int), st));
    LAUNCH(contact_elem_length(c, st));
    LAUNCH(contact_face_areas(c, st));
    ct.step_count = 0;
  }
  WFX_CHECK_HIP(hipStreamSynchronize(st));
  c->initialized = true;
  return scan_elements(c);
}

// ---- per-step API -----------------------------------------------------------
#define STEP_FN(name, call)                  \
  int name(wfx_ctx *c, wfx_stream s) {       \
    REQUIRE_MESH(c);                         \
    hipStream_t st = as_stream(s);           \
    (void)st;                                \
    LAUNCH(call);                            \
    return WFX_OK;                           \
  }

STEP_FN(wfx_update_prediction, launch_update_prediction(c, st))
STEP_FN(wfx_calc_elem_jacobians, BY_PRECISION(c, launch_elem_jacobians, c, st))
STEP_FN(wfx_calc_elem_vol, BY_PRECISION(c, launch_elem_vol, c, st))
int wfx_calc_elem_density(wfx_ctx *c, wfx_stream s) {
  REQUIRE_MESH(c);
  LAUNCH(BY_PRECISION(c, launch_elem_density, c, as_stream(s)));
  // unknown until the next scan: the fused pass loads densities and
  // evaluates the pressure terms in full
  c->rho_uniform = 0.0;
  c->elem_guard = 1;
  return WFX_OK;
}
STEP_FN(wfx_calc_nodal_vol, BY_PRECISION(c, launch_nodal_vol, c, st))
STEP_FN(wfx_calc_nodal_mass_from_vol, BY_PRECISION(c, launch_nodal_mass, c, st))
STEP_FN(wfx_calc_elem_strain_rates, BY_PRECISION(c, launch_strain_rates, c, st))
STEP_FN(wfx_calc_elem_pressure, BY_PRECISION(c, launch_pressure, c, st))
STEP_FN(wfx_calc_artificial_viscosity, BY_PRECISION(c, launch_artificial_viscosity, c, st))
STEP_FN(wfx_calc_elem_forces, BY_PRECISION(c, launch_elem_forces, c, st))
STEP_FN(wfx_calc_elem_hourglass_forces, BY_PRECISION(c, launch_hourglass, c, st))
STEP_FN(wfx_assembly_forces, BY_PRECISION(c, launch_assembly, c, st))
STEP_FN(wfx_calc_accel, launch_calc_accel(c, st))
STEP_FN(wfx_update_correction_accvel, launch_correction_accvel(c, st))
STEP_FN(wfx_update_correction_pos, launch_correction_pos(c, st))

int wfx_calc_stress_strain(wfx_ctx *c, double dt, wfx_stream s) {
  REQUIRE_MESH(c);
  LAUNCH(BY_PRECISION(c, launch_stress_strain, c, dt, as_stream(s)));
  return WFX_OK;
}

int wfx_impose_bcv(wfx_ctx *c, int dim, wfx_stream s) {
  REQUIRE_MESH(c);
  REQUIRE(dim >= 0 && dim < 3, WFX_ERR_ARG, "dim must be 0, 1 or 2");
  LAUNCH(launch_impose(c, dim, false, as_stream(s)));
  return WFX_OK;
}

int wfx_impose_bca(wfx_ctx *c, int dim, wfx_stream s) {
  REQUIRE_MESH(c);
  REQUIRE(dim >= 0 && dim < 3, WFX_ERR_ARG, "dim must be 0, 1 or 2");
  LAUNCH(launch_impose(c, dim, true, as_stream(s)));
  return WFX_OK;
}

// One SolveChungHulbert loop body (Solver_explicit.C:524-978, CPU order).
// When hist is given, the reaction force after assembly goes to *hist.
// per-step CFL time step (Solver_explicit.C:576-600) from the current
// positions and velocities; the per-step kernels take dt by value, so it
// comes back to the host
static int cfl_step(wfx_ctx *c, hipStream_t st) {
  LAUNCH(BY_PRECISION(c, launch_cfl_step, c, (int)c->dt_hist_h.size(), st));
  if (c->contact.on) LAUNCH(contact_elem_length(c, st));  // calcMinEdgeLength's element heights
  double dt = 0.0;
  WFX_CHECK_HIP(hipMemcpyAsync(&dt, c->scalars + kCflDt, sizeof(double), hipMemcpyDeviceToHost, st));
  WFX_CHECK_HIP(hipStreamSynchronize(st));
  c->prm.dt = dt;
  c->dt_hist_h.push_back(dt);
  return WFX_OK;
}

int wfx_calc_cfl_dt(wfx_ctx *c, double *dt, wfx_stream s) {
  REQUIRE_MESH(c);
  REQUIRE(c->initialized, WFX_ERR_STATE, "call wfx_init first");
  REQUIRE(c->prm.cfl > 0.0, WFX_ERR_STATE, "no CFL factor: call wfx_set_cfl");
  int rc = cfl_step(c, as_stream(s));
  if (rc) return rc;
  if (dt) *dt = c->prm.dt;
  return WFX_OK;
}

// End of a run: the device counters of non-finite internal forces (scrubbed
// to zero, Solver_explicit.C:777-785) and of elements with detJ <= 0 or not
// finite (Solver_explicit.C:639-652).  The run itself completes, as the
// reference's loop prints and goes on; WFX_ERR_NONFINITE reports that either
// counter grew during it.  Synchronises the stream (off with check_status 0).
static int check_run_status(wfx_ctx *c, hipStream_t st) {
  if (!c->check_status) return WFX_OK;
  unsigned int h[2] = {0, 0};
  WFX_CHECK_HIP(hipMemcpyAsync(h, c->status, sizeof(h), hipMemcpyDeviceToHost, st));
  WFX_CHECK_HIP(hipStreamSynchronize(st));
  const unsigned int nf = h[kStatNonfinite] - c->status_seen[0], inv = h[kStatInverted] - c->status_seen[1];
  c->status_seen[0] = h[kStatNonfinite];
  c->status_seen[1] = h[kStatInverted];
  if (nf == 0 && inv == 0) return WFX_OK;
  set_error("run completed with " + std::to_string(inv) + " element Jacobians with detJ <= 0 or not finite and " +
            std::to_string(nf) + " non-finite internal-force entries set to zero");
  return WFX_ERR_NONFINITE;
}

static int step_api(wfx_ctx *c, hipStream_t st, double *hist) {
  // external face areas every 10 steps, at the step's start (Solver_explicit.C:446-451)
  if (c->contact.on && c->contact.step_count % 10 == 0) LAUNCH(contact_face_areas(c, st));
  LAUNCH(launch_update_prediction(c, st));
  for (int d = 0; d < 3; ++d) LAUNCH(launch_impose(c, d, false, st));
  LAUNCH(BY_PRECISION(c, launch_elem_jacobians, c, st));
  if (c->prm.cfl > 0.0) {
    int rc = cfl_step(c, st);
    if (rc) return rc;
  }
  LAUNCH(BY_PRECISION(c, launch_elem_vol, c, st));
  LAUNCH(BY_PRECISION(c, launch_nodal_vol, c, st));
  LAUNCH(BY_PRECISION(c, launch_nodal_mass, c, st));
  LAUNCH(BY_PRECISION(c, launch_strain_rates, c, st));
  if (c->thermal.on) LAUNCH(BY_PRECISION(c, launch_thermal_expansion, c, st));  // Solver_explicit.C:719-720
  LAUNCH(BY_PRECISION(c, launch_pressure, c, st));
  LAUNCH(BY_PRECISION(c, launch_stress_strain, c, c->prm.dt, st));
  LAUNCH(BY_PRECISION(c, launch_artificial_viscosity, c, st));
  LAUNCH(BY_PRECISION(c, launch_elem_forces, c, st));
  LAUNCH(BY_PRECISION(c, launch_hourglass, c, st));
  if (c->contact.on) LAUNCH(contact_forces(c, st));  // CalcContactForces (Solver_explicit.C:769-770)
  LAUNCH(BY_PRECISION(c, launch_assembly, c, st));
  if (hist && c->ngroups) LAUNCH(launch_reaction(c, 2, hist, st));
  LAUNCH(launch_calc_accel(c, st));
  for (int d = 0; d < 3; ++d) LAUNCH(launch_impose(c, d, true, st));
  LAUNCH(launch_correction_accvel(c, st));
  for (int d = 0; d < 3; ++d) LAUNCH(launch_impose(c, d, false, st));
  LAUNCH(launch_correction_pos(c, st));
  if (c->contact.on) LAUNCH(contact_rigid_update(c, c->prm.dt, st));  // TriMesh_d::Move (:996-1004)
  if (c->thermal.on) LAUNCH(BY_PRECISION(c, launch_thermal, c, st));  // ThermalCalcs
  c->contact.step_count++;
  c->fused_stale = false;  // every per-step intermediate is current again
  return WFX_OK;
}

// device per-step dt record of a fused run with the per-step CFL time step
static int ensure_dt_hist(wfx_ctx *c, int nsteps) {
  if (c->prm.cfl > 0.0 && c->dt_cap < nsteps) {
    dfree(c, c->dt_hist);
    WFX_CHECK_HIP(dalloc(c, &c->dt_hist, nsteps));
    c->dt_cap = nsteps;
  }
  return WFX_OK;
}

// after a fused run with the per-step CFL time step: the host's dt and the
// run's dt record follow the device (synchronises)
static int sync_cfl_dt(wfx_ctx *c, int nsteps, hipStream_t st) {
  if (!(c->prm.cfl > 0.0)) return WFX_OK;
  c->dt_hist_h.assign(nsteps, 0.0);
  WFX_CHECK_HIP(hipMemcpyAsync(c->dt_hist_h.data(), c->dt_hist, nsteps * sizeof(double), hipMemcpyDeviceToHost, st));
  WFX_CHECK_HIP(hipMemcpyAsync(&c->prm.dt, c->scalars + kCflDt, sizeof(double), hipMemcpyDeviceToHost, st));
  WFX_CHECK_HIP(hipStreamSynchronize(st));
  return WFX_OK;
}

// device reaction history for nsteps steps of every reaction group
static int ensure_hist(wfx_ctx *c, int nsteps, double **out) {
  const long need = (long)nsteps * (c->ngroups > 0 ? c->ngroups : 1);
  if (c->hist_cap < need) {
    dfree(c, c->hist);
    WFX_CHECK_HIP(dalloc(c, &c->hist, need));
    c->hist_cap = (int)need;
  }
  *out = c->hist;
  return WFX_OK;
}

int wfx_step_api(wfx_ctx *c, wfx_stream s) {
  REQUIRE_MESH(c);
  REQUIRE(c->initialized, WFX_ERR_STATE, "call wfx_init first");
  return step_api(c, as_stream(s), nullptr);
}

int wfx_run(wfx_ctx *c, int nsteps, double *reaction_hist, wfx_stream s) {
  REQUIRE_MESH(c);
  REQUIRE(c->initialized, WFX_ERR_STATE, "call wfx_init first");
  REQUIRE(nsteps >= 0, WFX_ERR_ARG, "nsteps must be >= 0");
  hipStream_t st = as_stream(s);
  double *dh = nullptr;
  if (reaction_hist && nsteps > 0) {
    int rc = ensure_hist(c, nsteps, &dh);
    if (rc) return rc;
  }
  const int ng = c->ngroups > 0 ? c->ngroups : 1;
  c->dt_hist_h.clear();
  for (int k = 0; k < nsteps; ++k) {
    int rc = step_api(c, st, dh ? dh + (long)k * ng : nullptr);
    if (rc) return rc;
  }
  if (dh) {
    WFX_CHECK_HIP(hipMemcpyAsync(reaction_hist, dh, (size_t)nsteps * ng * sizeof(double), hipMemcpyDeviceToHost, st));
    WFX_CHECK_HIP(hipStreamSynchronize(st));
  }
  return check_run_status(c, st);
}

static int enqueue_fused(wfx_ctx *c, int nsteps, double *dev_hist, hipStream_t st);

int wfx_step_fused(wfx_ctx *c, int nsteps, double *reaction_hist, wfx_stream s) {
  REQUIRE_MESH(c);
  REQUIRE(c->initialized, WFX_ERR_STATE, "call wfx_init first");
  REQUIRE(nsteps >= 0, WFX_ERR_ARG, "nsteps must be >= 0");
  if (nsteps == 0) return WFX_OK;
  hipStream_t st = as_stream(s);
  double *dh = nullptr;
  if (reaction_hist) {
    int rc = ensure_hist(c, nsteps, &dh);
    if (rc) return rc;
  }
  int rc = ensure_dt_hist(c, nsteps);
  if (rc) return rc;
  // V0_n (vol0 may have been replaced through wfx_set_field), then the first
  // prediction; each fused step predicts the next one in its node pass.
  rc = enqueue_fused(c, nsteps, dh, st);
  if (rc) return rc;
  c->contact.step_count += nsteps;
  c->fused_stale = true;
  rc = sync_cfl_dt(c, nsteps, st);
  if (rc) return rc;
  if (dh) {
    const int ng = c->ngroups > 0 ? c->ngroups : 1;
    WFX_CHECK_HIP(hipMemcpyAsync(reaction_hist, dh, (size_t)nsteps * ng * sizeof(double), hipMemcpyDeviceToHost, st));
    WFX_CHECK_HIP(hipStreamSynchronize(st));
  }
  return check_run_status(c, st);
}

// ---- specimen batches, autotuning, graph capture --------------------------
int wfx_set_elem_tangent(wfx_ctx *c, const double *Et, int count) {
  REQUIRE_MESH(c);
  REQUIRE(Et && count == c->E, WFX_ERR_ARG, "Et needs one value per element");
  REQUIRE(c->prm.mat.model == WFX_BILINEAR && c->prm.mat.E > 0.0, WFX_ERR_STATE,
          "per-element tangent moduli need a bilinear material (wfx_set_material first)");
  const double E = c->prm.mat.E;
  std::vector<double> H(count);
  for (int k = 0; k < count; ++k) {
    REQUIRE(Et[k] >= 0.0 && Et[k] < E, WFX_ERR_ARG, "tangent modulus must be in [0, E)");
    H[k] = E * Et[k] / (E - Et[k]);
  }
  void **slot = c->precision == WFX_FP64 ? (void **)&c->e64.He : (void **)&c->e32.He;
  if (!*slot) {
    if (c->precision == WFX_FP64) WFX_CHECK_HIP(dalloc(c, &c->e64.He, c->E));
    else WFX_CHECK_HIP(dalloc(c, &c->e32.He, c->E));
  }
  drop_graph(c);  // kernel arguments changed
  return upload_elem_values(c, *slot, H.data(), (long)H.size());
}

}  // extern "C"

// Buffers the fused phases write; saved and restored around autotuning.
std::vector<std::pair<void *, size_t>> wfx::fused_state_buffers(wfx_ctx *c) {
  const size_t N3 = 3UL * c->N * sizeof(double), N1 = (size_t)c->N * sizeof(double);
  const size_t es = c->precision == WFX_FP64 ? sizeof(double) : sizeof(float), E = c->E;
  std::vector<std::pair<void *, size_t>> v = {
      {c->nd.x, N3}, {c->nd.v, N3}, {c->nd.u, N3}, {c->nd.u_dt, N3}, {c->nd.prev_a, N3}, {c->nd.fi, N3},
      {c->nd.mdiag, N1}, {c->nd.voln, N1}, {c->nd.anp_ratio, N1}, {c->status, 16 * sizeof(unsigned)},
      {c->nd.T, N1}};
  if (c->contact.on) {  // the contact kernel runs in the element phase
    const auto &ct = c->contact;
    v.push_back({ct.contforce, N3});
    v.push_back({ct.ut_prev, N3});
    v.push_back({ct.cfn, N1});
    v.push_back({ct.mesh_in_contact, (size_t)c->N * sizeof(int)});
  }
  auto add = [&](void *p, size_t rows) {
    if (p) v.push_back({p, rows * E * es});
  };
  if (c->precision == WFX_FP64) {
    auto &el = c->e64;
    add(el.vol, 1); add(el.tau, 6); add(el.eps, 6); add(el.p, 1); add(el.pl_strain, 1); add(el.sigma_y, 1);
    add(el.q_plheat, 1);
  } else {
    auto &el = c->e32;
    add(el.vol, 1); add(el.tau, 6); add(el.eps, 6); add(el.p, 1); add(el.pl_strain, 1); add(el.sigma_y, 1);
    add(el.q_plheat, 1);
  }
  return v;
}

extern "C" {

int wfx_autotune(wfx_ctx *c, int reps, int *chosen, wfx_stream s) {
  REQUIRE_MESH(c);
  REQUIRE(c->initialized, WFX_ERR_STATE, "call wfx_init first");
  REQUIRE(reps > 0, WFX_ERR_ARG, "reps must be positive");
  hipStream_t st = as_stream(s);
  auto state = fused_state_buffers(c);
  // every failure below still reaches the restore and the frees at the end
  hipError_t err = hipSuccess;
  std::vector<void *> saved(state.size(), nullptr);
  for (size_t k = 0; k < state.size() && err == hipSuccess; ++k) {
    err = hipMalloc(&saved[k], state[k].second);
    if (err == hipSuccess)
      err = hipMemcpyAsync(saved[k], state[k].first, state[k].second, hipMemcpyDeviceToDevice, st);
  }
  const bool have_copy = err == hipSuccess;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (err == hipSuccess) err = hipEventCreate(&e0);
  if (err == hipSuccess) err = hipEventCreate(&e1);
  static const int sizes[4] = {64, 128, 256, 512};
  // whole fused steps (the four passes in order), so that each candidate sees
  // the caches its predecessor in the step leaves behind
  auto step = [&]() {
    for (int ph = 1; ph <= 4 && err == hipSuccess; ++ph) err = BY_PRECISION(c, launch_fused_phase, c, ph, 1, st, -1);
  };
  // coordinate descent over the four passes, two sweeps; within a pass five
  // rounds over the candidates of at least 10 steps each, the best round of
  // each counts: interleaving the candidates keeps clock and cache drift out
  // of the choice (with three rounds of 5 steps a pass of 0.66 ms was once
  // tuned to a choice 4% slower)
  const int nrep = reps > 10 ? reps : 10;
  // the element passes (phases 1 and 3) run one work-group per element cluster
  if (chosen) chosen[0] = chosen[2] = kCluster;
  for (int sweep = 0; sweep < 2 && err == hipSuccess; ++sweep)
    for (int ph = 2; ph <= 4 && err == hipSuccess; ++ph) {
      if (ph == 3) continue;
      float best_t[4] = {1e30f, 1e30f, 1e30f, 1e30f};
      for (int round = 0; round < 5 && err == hipSuccess; ++round)
        for (int k = 0; k < 4 && err == hipSuccess; ++k) {
          c->fused_bs[ph] = sizes[k];
          step();  // warm-up
          if (err == hipSuccess) err = hipEventRecord(e0, st);
          for (int r = 0; r < nrep && err == hipSuccess; ++r) step();
          if (err == hipSuccess) err = hipEventRecord(e1, st);
          if (err == hipSuccess) err = hipEventSynchronize(e1);
          float ms = 0.f;
          if (err == hipSuccess) err = hipEventElapsedTime(&ms, e0, e1);
          if (err == hipSuccess && ms < best_t[k]) best_t[k] = ms;
        }
      int best_bs = 256;
      float best = 1e30f;
      for (int k = 0; k < 4; ++k)
        if (best_t[k] < best) { best = best_t[k]; best_bs = sizes[k]; }
      c->fused_bs[ph] = best_bs;
      if (chosen) chosen[ph - 1] = best_bs;
    }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  // the state is put back whenever the copy was taken, also after a failed
  // measurement (the launches may have advanced it)
  if (have_copy) {
    hipError_t r = hipSuccess;
    for (size_t k = 0; k < state.size() && r == hipSuccess; ++k)
      r = hipMemcpyAsync(state[k].first, saved[k], state[k].second, hipMemcpyDeviceToDevice, st);
    if (r == hipSuccess) r = hipStreamSynchronize(st);
    if (err == hipSuccess) err = r;
  } else {
    (void)hipStreamSynchronize(st);
  }
  for (void *p : saved)
    if (p) (void)hipFree(p);
  drop_graph(c);
  WFX_CHECK_HIP(err);
  return WFX_OK;
}

int wfx_set_block_sizes(wfx_ctx *c, const int *bs4) {
  REQUIRE_CTX(c);
  REQUIRE(bs4, WFX_ERR_ARG, "null block sizes");
  for (int k = 0; k < 4; ++k)
    REQUIRE(bs4[k] == 64 || bs4[k] == 128 || bs4[k] == 256 || bs4[k] == 512, WFX_ERR_ARG,
            "block sizes must be 64, 128, 256 or 512");
  for (int k = 0; k < 4; ++k) c->fused_bs[k + 1] = bs4[k];
  drop_graph(c);
  return WFX_OK;
}

// Everything a captured graph bakes into its kernel arguments; a launch
// refuses to run when it changed since the capture.
static uint64_t graph_signature(wfx_ctx *c) {
  uint64_t h = 1469598103934665603ULL;
  auto mix = [&](const void *p, size_t n) {
    const unsigned char *b = (const unsigned char *)p;
    for (size_t k = 0; k < n; ++k) h = (h ^ b[k]) * 1099511628211ULL;
  };
  // with the per-step CFL time step the graph reads dt from the device
  const double dts[5] = {c->prm.cfl > 0.0 ? 0.0 : c->prm.dt, c->prm.alpha, c->prm.beta, c->prm.gamma,
                         c->prm.cfl};
  mix(dts, sizeof(dts));
  mix(&c->prm.mat, sizeof(c->prm.mat));
  mix(&c->prm.stab, sizeof(c->prm.stab));
  mix(&c->thermal, sizeof(c->thermal));
  {
    const auto &ct = c->contact;
    const double cp[5] = {ct.pf, ct.mu_sta, ct.mu_dyn, ct.heat_cond, ct.T_const};
    const int ci[4] = {ct.on, ct.rn, ct.rt, ct.next};
    const void *cptr[4] = {ct.rnode, ct.contforce, ct.ext_nodes, ct.q_cont};
    mix(cp, sizeof(cp));
    mix(ci, sizeof(ci));
    mix(cptr, sizeof(cptr));
  }
  const void *ptrs[] = {c->bc_nod[0], c->bc_nod[1], c->bc_nod[2], c->bc_slot, c->bc_mask, c->bc_tab,
                        c->react, c->hist, c->e64.He, c->e32.He, c->temp_e, c->dt_hist};
  mix(ptrs, sizeof(ptrs));
  int ints[] = {c->N, c->E, c->n, c->has_fe, c->bc_count[0], c->bc_count[1], c->bc_count[2],
                (int)c->react_h.size(), c->fused_bs[1], c->fused_bs[2], c->fused_bs[3], c->fused_bs[4]};
  mix(ints, sizeof(ints));
  mix(&c->rho_uniform, sizeof(double));
  mix(&c->elem_guard, sizeof(int));
  mix(&c->rho_min, sizeof(double));
  return h;
}

// Enqueue the body of wfx_step_fused (kernels only: it is what a graph holds).
static int enqueue_fused(wfx_ctx *c, int nsteps, double *dev_hist, hipStream_t st) {
  LAUNCH(BY_PRECISION(c, launch_vsum0, c, st));
  LAUNCH(launch_update_prediction(c, st));
  for (int d = 0; d < 3; ++d) LAUNCH(launch_impose(c, d, false, st));
  LAUNCH(BY_PRECISION(c, launch_fused_steps, c, nsteps, dev_hist, st, (hipEvent_t *)nullptr));
  LAUNCH(launch_copy(c->nd.prev_a, c->nd.a, 3L * c->N, st));  // a == prev_a after the step
  return WFX_OK;
}

// The fused step of wfx_step_fused with HIP events recorded on the stream
// before each of its four passes and after the last step: the average
// duration of each pass in microseconds over nsteps steps, measured in the
// order the passes run (each sees the caches its predecessor leaves).
int wfx_time_fused_phases(wfx_ctx *c, int nsteps, double *us4, wfx_stream s) {
  REQUIRE_MESH(c);
  REQUIRE(c->initialized, WFX_ERR_STATE, "call wfx_init first");
  REQUIRE(nsteps > 0 && us4, WFX_ERR_ARG, "nsteps must be positive and us4 non-null");
  int rc = ensure_dt_hist(c, nsteps);
  if (rc) return rc;
  hipStream_t st = as_stream(s);
  std::vector<hipEvent_t> ev(4 * (size_t)nsteps + 1, nullptr);
  hipError_t e = hipSuccess;
  for (auto &x : ev)
    if (e == hipSuccess) e = hipEventCreate(&x);
  if (e == hipSuccess) e = BY_PRECISION(c, launch_vsum0, c, st);
  if (e == hipSuccess) e = launch_update_prediction(c, st);
  for (int d = 0; d < 3 && e == hipSuccess; ++d) e = launch_impose(c, d, false, st);
  if (e == hipSuccess) e = BY_PRECISION(c, launch_fused_steps, c, nsteps, nullptr, st, ev.data());
  if (e == hipSuccess) e = launch_copy(c->nd.prev_a, c->nd.a, 3L * c->N, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  double acc[4] = {0, 0, 0, 0};
  for (int k = 0; k < nsteps && e == hipSuccess; ++k)
    for (int ph = 0; ph < 4 && e == hipSuccess; ++ph) {
      float ms = 0.f;
      e = hipEventElapsedTime(&ms, ev[4 * k + ph], ev[4 * k + ph + 1]);
      acc[ph] += ms;
    }
  for (auto &x : ev)
    if (x) (void)hipEventDestroy(x);
  WFX_CHECK_HIP(e);
  for (int ph = 0; ph < 4; ++ph) us4[ph] = acc[ph] * 1e3 / nsteps;
  c->contact.step_count += nsteps;
  c->fused_stale = true;
  rc = sync_cfl_dt(c, nsteps, st);
  return rc ? rc : check_run_status(c, st);
}

int wfx_graph_capture(wfx_ctx *c, int nsteps, int with_reaction, wfx_stream s) {
  REQUIRE_MESH(c);
  REQUIRE(c->initialized, WFX_ERR_STATE, "call wfx_init first");
  REQUIRE(nsteps > 0, WFX_ERR_ARG, "nsteps must be positive");
  REQUIRE(s != nullptr, WFX_ERR_ARG, "graph capture needs a created stream (not the null stream)");
  hipStream_t st = as_stream(s);
  drop_graph(c);
  double *dh = nullptr;
  if (with_reaction) {
    int rc = ensure_hist(c, nsteps, &dh);
    if (rc) return rc;
  }
  {
    int rc = ensure_dt_hist(c, nsteps);
    if (rc) return rc;
  }
  WFX_CHECK_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
  int rc = enqueue_fused(c, nsteps, dh, st);
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(st, &g);
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  WFX_CHECK_HIP(e);
  c->graph = g;
  WFX_CHECK_HIP(hipGraphInstantiate(&c->graph_exec, g, nullptr, nullptr, 0));
  c->graph_steps = nsteps;
  c->graph_hist = dh != nullptr;
  c->graph_phase = (int)(c->contact.step_count % 10);
  c->graph_sig = graph_signature(c);
  return WFX_OK;
}

int wfx_graph_launch(wfx_ctx *c, double *reaction_hist, wfx_stream s) {
  REQUIRE_MESH(c);
  REQUIRE(c->graph_exec, WFX_ERR_STATE, "no captured graph: call wfx_graph_capture");
  REQUIRE(c->graph_sig == graph_signature(c), WFX_ERR_STATE,
          "parameters, boundary conditions or block sizes changed since wfx_graph_capture: capture again");
  REQUIRE(!reaction_hist || c->graph_hist, WFX_ERR_STATE, "graph was captured without the reaction history");
  REQUIRE(!c->contact.on || c->contact.step_count % 10 == c->graph_phase, WFX_ERR_STATE,
          "with contact the graph's face-area refresh follows the step counter mod 10: capture a "
          "multiple of 10 steps, or capture again");
  hipStream_t st = as_stream(s);
  WFX_CHECK_HIP(hipGraphLaunch(c->graph_exec, st));
  c->contact.step_count += c->graph_steps;
  c->fused_stale = true;
  if (reaction_hist) {
    const int ng = c->ngroups > 0 ? c->ngroups : 1;
    WFX_CHECK_HIP(hipMemcpyAsync(reaction_hist, c->hist, (size_t)c->graph_steps * ng * sizeof(double),
                                 hipMemcpyDeviceToHost, st));
    WFX_CHECK_HIP(hipStreamSynchronize(st));
  }
  int rc = sync_cfl_dt(c, c->graph_steps, st);
  return rc ? rc : check_run_status(c, st);
}

int wfx_fused_phase(wfx_ctx *c, int phase, int predict, wfx_stream s) {
  REQUIRE_MESH(c);
  REQUIRE(c->initialized, WFX_ERR_STATE, "call wfx_init first");
  REQUIRE(phase >= 1 && phase <= 5, WFX_ERR_ARG, "phase must be 1..5");
  LAUNCH(BY_PRECISION(c, launch_fused_phase, c, phase, predict, as_stream(s), -1));
  c->fused_stale = true;
  return WFX_OK;
}

int wfx_reaction_force(wfx_ctx *c, int dim, double *out, wfx_stream s) {
  REQUIRE_MESH(c);
  REQUIRE(out && dim >= 0 && dim < 3, WFX_ERR_ARG, "bad arguments");
  REQUIRE(c->ngroups > 0, WFX_ERR_STATE, "no reaction nodes: call wfx_set_reaction_nodes");
  hipStream_t st = as_stream(s);
  double *dh = nullptr;
  int rc = ensure_hist(c, 1, &dh);
  if (rc) return rc;
  LAUNCH(launch_reaction(c, dim, dh, st));
  WFX_CHECK_HIP(hipMemcpyAsync(out, dh, c->ngroups * sizeof(double), hipMemcpyDeviceToHost, st));
  WFX_CHECK_HIP(hipStreamSynchronize(st));
  return WFX_OK;
}

// computeEnergies with the reduction partials in `partial` (2 doubles per
// 256-thread block over max(N, E), every one written before it is read)
static hipError_t energies_into(wfx_ctx *c, double *partial, double *ekin, double *deint, hipStream_t st) {
  hipError_t e = BY_PRECISION(c, launch_energies, c, c->scalars + 2, partial, st);
  double h[2] = {0, 0};
  if (e == hipSuccess) e = hipMemcpyAsync(h, c->scalars + 2, 2 * sizeof(double), hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e == hipSuccess) {
    *ekin = h[0];
    *deint = h[1] * c->prm.dt;
  }
  return e;
}

int wfx_energies(wfx_ctx *c, double *ekin, double *deint, wfx_stream s) {
  REQUIRE_MESH(c);
  REQUIRE(ekin && deint, WFX_ERR_ARG, "null output");
  REQUIRE(!c->fused_stale, WFX_ERR_STATE, "energies need the strain rates of a per-step API step");
  hipStream_t st = as_stream(s);
  int nb = (std::max(c->N, c->E) + 255) / 256;
  double *partial = nullptr;
  WFX_CHECK_HIP(hipMalloc(&partial, sizeof(double) * 2 * (nb + 1)));
  hipError_t e = energies_into(c, partial, ekin, deint, st);
  (void)hipFree(partial);
  WFX_CHECK_HIP(e);
  return WFX_OK;
}

// wfx_energies without allocating: the partials go to this device's copy of
// the static scratch array.  Calls that overlap on one GPU share it and race.
int wfx_energies_noalloc(wfx_ctx *c, double *ekin, double *deint, wfx_stream s) {
  REQUIRE_MESH(c);
  REQUIRE(ekin && deint, WFX_ERR_ARG, "null output");
  REQUIRE(!c->fused_stale, WFX_ERR_STATE, "energies need the strain rates of a per-step API step");
  REQUIRE(c->scratch, WFX_ERR_STATE, "no static scratch address for this context's device");
  const long nb = (std::max<long>(c->N, c->E) + 255) / 256;
  REQUIRE(2 * nb <= kScratchDoubles, WFX_ERR_ARG,
          "wfx_energies_noalloc: max(node count, element count) = " + std::to_string(std::max(c->N, c->E)) +
              " exceeds the static scratch (at most " + std::to_string(kScratchDoubles / 2 * 256) +
              "); use wfx_energies");
  WFX_CHECK_HIP(energies_into(c, c->scratch, ekin, deint, as_stream(s)));
  return WFX_OK;
}

int wfx_nonfinite_count(wfx_ctx *c, long *out) {
  REQUIRE_MESH(c);
  REQUIRE(out, WFX_ERR_ARG, "null output");
  unsigned int h = 0;
  WFX_CHECK_HIP(hipMemcpy(&h, c->status + kStatNonfinite, sizeof(unsigned int), hipMemcpyDeviceToHost));
  *out = h;
  return WFX_OK;
}

int wfx_inverted_count(wfx_ctx *c, long *out) {
  REQUIRE_MESH(c);
  REQUIRE(out, WFX_ERR_ARG, "null output");
  unsigned int h = 0;
  WFX_CHECK_HIP(hipMemcpy(&h, c->status + kStatInverted, sizeof(unsigned int), hipMemcpyDeviceToHost));
  *out = h;
  return WFX_OK;
}

// ---- bandwidth probe ----------------------------------------------------------
// A streaming copy in the access pattern that reaches the highest rate measured
// on MI355X: one 16-byte pair per thread, non-temporal loads and stores,
// 256-thread blocks (6.5 TB/s read + write on 2 GiB, where an 8-byte copy in
// 64-thread blocks reaches 4.8).  The benchmark's ceiling for HBM-bound passes.
typedef double wfx_d2v __attribute__((ext_vector_type(2)));
__global__ __launch_bounds__(256) void k_stream_copy(const wfx_d2v *__restrict__ a, wfx_d2v *__restrict__ b, long n) {
  const long i = blockIdx.x * 256L + threadIdx.x;
  if (i < n) __builtin_nontemporal_store(__builtin_nontemporal_load(&a[i]), &b[i]);
}

int wfx_stream_copy(const void *src, void *dst, long bytes, wfx_stream s) {
  REQUIRE(src && dst && bytes > 0 && bytes % 16 == 0 && ((uintptr_t)src | (uintptr_t)dst) % 16 == 0, WFX_ERR_ARG,
          "stream copy needs 16-byte aligned pointers and a positive multiple of 16 bytes");
  const long n = bytes / 16;
  REQUIRE((n + 255) / 256 < 2147483647L, WFX_ERR_ARG, "stream copy too large for one launch");
  hipLaunchKernelGGL(k_stream_copy, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, as_stream(s),
                     (const wfx_d2v *)src, (wfx_d2v *)dst, n);
  WFX_CHECK_HIP(hipGetLastError());
  return WFX_OK;
}

// ---- state transfer ---------------------------------------------------------
int wfx_field_count(wfx_ctx *c, const char *name, long *count, int *is_int) {
  REQUIRE_MESH(c);
  FieldDesc f;
  REQUIRE(name && find_field(c, name, f), WFX_ERR_ARG, std::string("unknown field ") + (name ? name : ""));
  if (count) *count = f.cnt * f.rows;
  if (is_int) *is_int = f.is_int;
  return WFX_OK;
}

// Per-step intermediates that the fused path keeps in registers only.
static bool fused_intermediate(const char *name) {
  static const char *k[] = {"detJ", "dHdx", "dHdy", "dHdz", "str_rate", "rot_rate",
                            "strain_pl_incr", "f_elem", "f_hg"};
  for (const char *s : k)
    if (!std::strcmp(s, name)) return true;
  return false;
}

int wfx_get_field(wfx_ctx *c, const char *name, void *host, long count) {
  REQUIRE_MESH(c);
  FieldDesc f;
  REQUIRE(name && find_field(c, name, f), WFX_ERR_ARG, std::string("unknown field ") + (name ? name : ""));
  REQUIRE(host && count == f.cnt * f.rows, WFX_ERR_ARG, std::string("size mismatch for field ") + name);
  REQUIRE(!(c->fused_stale && fused_intermediate(name)), WFX_ERR_STATE,
          std::string(name) + " is a per-step intermediate not kept by wfx_step_fused; "
          "call the per-step entry point that produces it first");
  WFX_CHECK_HIP(hipDeviceSynchronize());
  const long cnt = f.cnt;
  // device column of host item i (element fields are stored in cluster order)
  auto col = [&](long i) { return f.per_elem ? (long)c->el_o2s[i] : i; };
  if (f.is_int) {
    std::vector<int> tmp(cnt * f.rows);
    WFX_CHECK_HIP(hipMemcpy(tmp.data(), f.dev, tmp.size() * sizeof(int), hipMemcpyDeviceToHost));
    int *h = (int *)host;
    for (long i = 0; i < cnt; ++i)
      for (int r = 0; r < f.rows; ++r) {
        const int v = tmp[(long)r * cnt + col(i)];
        h[i * f.rows + r] = f.elem_ids ? c->el_s2o[v] : v;
      }
    return WFX_OK;
  }
  int drows = f.kind == 1 ? 3 : f.rows;
  std::vector<double> tmp((size_t)drows * cnt);
  int rc = download_real(c, tmp.data(), f.dev, tmp.size(), f.elem);
  if (rc) return rc;
  double *h = (double *)host;
  if (f.kind == 1) {
    for (long i = 0; i < cnt; ++i) {
      h[i * 6 + 0] = h[i * 6 + 1] = h[i * 6 + 2] = 0.0;
      for (int r = 0; r < 3; ++r) h[i * 6 + 3 + r] = tmp[(long)r * cnt + col(i)];
    }
  } else {
    for (long i = 0; i < cnt; ++i)
      for (int r = 0; r < f.rows; ++r) h[i * f.rows + r] = tmp[(long)r * cnt + col(i)];
  }
  return WFX_OK;
}

int wfx_set_field(wfx_ctx *c, const char *name, const void *host, long count) {
  REQUIRE_MESH(c);
  FieldDesc f;
  REQUIRE(name && find_field(c, name, f), WFX_ERR_ARG, std::string("unknown field ") + (name ? name : ""));
  REQUIRE(host && count == f.cnt * f.rows, WFX_ERR_ARG, std::string("size mismatch for field ") + name);
  REQUIRE(!f.is_int, WFX_ERR_ARG, "connectivity and CSR are set through the mesh calls");
  WFX_CHECK_HIP(hipDeviceSynchronize());
  const long cnt = f.cnt;
  const double *h = (const double *)host;
  int drows = f.kind == 1 ? 3 : f.rows;
  std::vector<double> tmp((size_t)drows * cnt);
  auto col = [&](long i) { return f.per_elem ? (long)c->el_o2s[i] : i; };
  if (f.kind == 1) {
    for (long i = 0; i < cnt; ++i)
      for (int r = 0; r < 3; ++r) tmp[(long)r * cnt + col(i)] = h[i * 6 + 3 + r];
  } else {
    for (long i = 0; i < cnt; ++i)
      for (int r = 0; r < f.rows; ++r) tmp[(long)r * cnt + col(i)] = h[i * f.rows + r];
  }
  if (!std::strcmp(name, "fe")) {
    c->has_fe = 0;
    for (double v : tmp)
      if (v != 0.0) { c->has_fe = 1; break; }
  }
  int rc = upload_real(c, f.dev, tmp.data(), tmp.size(), f.elem);
  if (rc == WFX_OK && (!std::strcmp(name, "rho") || !std::strcmp(name, "vol0"))) rc = scan_elements(c);
  return rc;
}

}  // extern "C"
