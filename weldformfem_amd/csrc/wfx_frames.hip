// Asynchronous VTK frames (the "async device-to-host frames" beside the
// step).  A frame is a snapshot of the state fields: device-to-device copies
// into a staging area on the caller's stream (ordered with the step kernels,
// at HBM speed), then one device-to-host copy into pinned memory on a copy
// stream of the library, then a host thread that waits for that copy,
// converts the arrays to the reference's layout and writes the file.  The
// step loop continues as soon as the staging copies are queued; two frames
// can be in flight.
#include <hip/hip_runtime.h>

#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../include/wfx.h"
#include "wfx_internal.hpp"
#include "wfx_vtk.hpp"

#define REQUIRE(cond, code, msg) \
  do {                           \
    if (!(cond)) {               \
      wfx::set_error(msg);       \
      return code;               \
    }                            \
  } while (0)
#define REQUIRE_CTX(c) REQUIRE((c) != nullptr, WFX_ERR_ARG, "null context")
#define REQUIRE_MESH(c) \
  REQUIRE_CTX(c);       \
  REQUIRE((c)->has_mesh, WFX_ERR_STATE, "no mesh: call wfx_add_box_length or wfx_set_mesh first")

namespace wfx {

struct FrameSlot {
  void *dev = nullptr;   // staging on the device
  void *host = nullptr;  // pinned host copy
  size_t bytes = 0;
  hipEvent_t snap = nullptr, done = nullptr;
  std::thread writer;
  int status = 0;
  std::string err;
};

struct Frames {
  hipStream_t copy = nullptr;
  FrameSlot slot[2];
  int next = 0;
  int status = 0;  // first failure of a finished frame, reported by wfx_vtk_wait
  std::string err;
};

static void join_slot(Frames *fr, FrameSlot &s) {
  if (s.writer.joinable()) s.writer.join();
  if (s.status && !fr->status) {
    fr->status = s.status;
    fr->err = s.err;
  }
  s.status = 0;
}

// Waits for the frames in flight (their writer threads read the host
// connectivity); a failure stays recorded for wfx_vtk_wait.
void frames_join(wfx_ctx *c) {
  Frames *fr = c->frames;
  if (!fr) return;
  for (auto &s : fr->slot) join_slot(fr, s);
}

void frames_free(wfx_ctx *c) {
  Frames *fr = c->frames;
  if (!fr) return;
  for (auto &s : fr->slot) {
    if (s.writer.joinable()) s.writer.join();
    if (s.dev) (void)hipFree(s.dev);
    if (s.host) (void)hipHostFree(s.host);
    if (s.snap) (void)hipEventDestroy(s.snap);
    if (s.done) (void)hipEventDestroy(s.done);
  }
  if (fr->copy) (void)hipStreamDestroy(fr->copy);
  delete fr;
  c->frames = nullptr;
}

}  // namespace wfx

using namespace wfx;

extern "C" {

int wfx_write_vtk_async(wfx_ctx *c, const char *path, wfx_stream s) {
  REQUIRE_MESH(c);
  REQUIRE(path && *path, WFX_ERR_ARG, "null path");
  if (!c->frames) {
    c->frames = new Frames();
    WFX_CHECK_HIP(hipStreamCreate(&c->frames->copy));  // default priority
  }
  Frames *fr = c->frames;
  FrameSlot &sl = fr->slot[fr->next];
  fr->next ^= 1;
  join_slot(fr, sl);  // the slot's previous frame has been written
  const size_t N = c->N, E = c->E, n = c->n;
  const size_t es = c->precision == WFX_FP64 ? sizeof(double) : sizeof(float);
  const size_t nb = 13 * N * sizeof(double), bytes = nb + 10 * E * es;
  if (sl.bytes != bytes) {
    if (sl.dev) WFX_CHECK_HIP(hipFree(sl.dev));
    if (sl.host) WFX_CHECK_HIP(hipHostFree(sl.host));
    sl.dev = sl.host = nullptr;
    sl.bytes = 0;
    WFX_CHECK_HIP(hipMalloc(&sl.dev, bytes));
    WFX_CHECK_HIP(hipHostMalloc(&sl.host, bytes, hipHostMallocDefault));
    sl.bytes = bytes;
  }
  if (!sl.snap) WFX_CHECK_HIP(hipEventCreateWithFlags(&sl.snap, hipEventDisableTiming));
  if (!sl.done) WFX_CHECK_HIP(hipEventCreateWithFlags(&sl.done, hipEventDisableTiming));
  hipStream_t st = (hipStream_t)s;
  // staging layout: x u v a [3][N], mdiag [N] (double); p pl_strain sigma_y
  // vol [E], sigma [6][E] (element real type)
  char *d = (char *)sl.dev;
  const double *nsrc[4] = {c->nd.x, c->nd.u, c->nd.v, c->nd.a};
  for (int k = 0; k < 4; ++k)
    WFX_CHECK_HIP(hipMemcpyAsync(d + k * 3 * N * sizeof(double), nsrc[k], 3 * N * sizeof(double),
                                 hipMemcpyDeviceToDevice, st));
  WFX_CHECK_HIP(hipMemcpyAsync(d + 12 * N * sizeof(double), c->nd.mdiag, N * sizeof(double),
                               hipMemcpyDeviceToDevice, st));
  const void *esrc[5];
  if (c->precision == WFX_FP64) {
    const auto &el = c->e64;
    const void *t[5] = {el.p, el.pl_strain, el.sigma_y, el.vol, el.sigma};
    std::memcpy(esrc, t, sizeof(t));
  } else {
    const auto &el = c->e32;
    const void *t[5] = {el.p, el.pl_strain, el.sigma_y, el.vol, el.sigma};
    std::memcpy(esrc, t, sizeof(t));
  }
  for (int k = 0; k < 5; ++k)
    WFX_CHECK_HIP(hipMemcpyAsync(d + nb + k * E * es, esrc[k], (k == 4 ? 6 : 1) * E * es,
                                 hipMemcpyDeviceToDevice, st));
  WFX_CHECK_HIP(hipEventRecord(sl.snap, st));
  WFX_CHECK_HIP(hipStreamWaitEvent(fr->copy, sl.snap, 0));
  WFX_CHECK_HIP(hipMemcpyAsync(sl.host, sl.dev, bytes, hipMemcpyDeviceToHost, fr->copy));
  WFX_CHECK_HIP(hipEventRecord(sl.done, fr->copy));
  const bool f64 = c->precision == WFX_FP64;
  const std::vector<int> *elnod = &c->elnod_h;  // stable until the mesh changes (which joins first)
  const std::vector<int> *o2s = &c->el_o2s;     // element fields are stored in cluster order
  std::string p(path);
  FrameSlot *slp = &sl;
  sl.writer = std::thread([slp, N, E, n, nb, es, f64, elnod, o2s, p]() {
    if (hipEventSynchronize(slp->done) != hipSuccess) {
      slp->status = WFX_ERR_HIP;
      slp->err = "frame copy failed";
      return;
    }
    const double *h = (const double *)slp->host;
    VtkFrame fr;
    fr.N = (int)N;
    fr.E = (int)E;
    fr.n = (int)n;
    std::vector<double> *nd[4] = {&fr.x, &fr.u, &fr.v, &fr.a};
    for (int k = 0; k < 4; ++k) {
      nd[k]->resize(3 * N);
      const double *q = h + k * 3 * N;
      for (size_t i = 0; i < N; ++i)
        for (int r = 0; r < 3; ++r) (*nd[k])[3 * i + r] = q[r * N + i];
    }
    fr.mdiag.assign(h + 12 * N, h + 13 * N);
    auto elv = [&](size_t row) {
      const char *b = (const char *)slp->host + nb;
      return f64 ? ((const double *)b)[row] : (double)((const float *)b)[row];
    };
    std::vector<double> *ef[4] = {&fr.p, &fr.pl_strain, &fr.sigma_y, &fr.vol};
    for (int k = 0; k < 4; ++k) {
      ef[k]->resize(E);
      for (size_t e = 0; e < E; ++e) (*ef[k])[e] = elv(k * E + (*o2s)[e]);
    }
    fr.sigma.resize(6 * E);
    for (size_t e = 0; e < E; ++e)
      for (int r = 0; r < 6; ++r) fr.sigma[6 * e + r] = elv(4 * E + r * E + (*o2s)[e]);
    fr.elnod = *elnod;
    std::string err;
    if (write_vtk_frame(p.c_str(), fr, err)) {
      slp->status = WFX_ERR_ARG;
      slp->err = err;
    }
  });
  return WFX_OK;
}

int wfx_vtk_wait(wfx_ctx *c) {
  REQUIRE_CTX(c);
  Frames *fr = c->frames;
  if (!fr) return WFX_OK;
  for (auto &s : fr->slot) join_slot(fr, s);
  const int st = fr->status;
  if (st) {
    set_error(fr->err);
    fr->status = 0;
    fr->err.clear();
  }
  return st;
}

}  // extern "C"
