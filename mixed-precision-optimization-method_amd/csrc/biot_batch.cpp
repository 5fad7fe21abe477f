// Batched shots: many windows of one model advanced together (include/biot.h, biot_batch_*).
//
// Every shot is a grid of its own (a biot_handle: its window of the model, its absorbing
// boundary, source and receivers).  One persistent launch per half step walks the tiles of
// all shots (k_step_batch), the per-shot argument tables live in device memory, and the step
// index comes from a device counter, so a run of steps is captured once into a HIP graph and
// replayed.  The gradient runs the checkpointed forward, the FP64 misfit, the segment
// recompute and the adjoint sweep for all shots at once and adds every shot's parameter
// gradient into one model-grid volume on the device.  The reference has no counterpart.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/biot.h"
#include "biot_adjoint.hpp"
#include "biot_handle.hpp"

using namespace biot;

struct biot_batch {
  biot_batch_config cfg{};
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  std::string err;
  double last_ms = 0;

  struct Shot {
    biot_handle *h = nullptr;
    int x0, y0, z0;
    std::vector<void *> ckpt;      // checkpoint states
    double *d_par[8] = {};         // window parameters (chain rule)
    float *d_obs = nullptr;        // observed traces
  };
  std::vector<Shot> shots;

  LaunchShape shape{64, 4, 8};
  int blocks = 2048;               // persistent grid of the batched stencil launches

  // forward tables
  StepArgs *d_step = nullptr;
  int *d_tiles = nullptr;
  SourceArgs *d_src = nullptr;
  RecordArgs *d_rec = nullptr;
  int *d_it = nullptr;
  int *d_queue = nullptr;          // 16 ints, the persistent grid's work-queue counters
  int total_tiles = 0;
  int nactive = 0;                 // shots the current tables cover
  int64_t max_nrec = 0;
  int max_ncomp = 0;
  bool built = false;
  // captured steps
  hipGraphExec_t gexec = nullptr;
  int g_steps = 0;
  bool g_record = false;
  // gradient tables
  AdjArgs *d_adj = nullptr;
  int *d_blocks = nullptr;
  InjectArgs *d_inj = nullptr;
  CopyEntry *d_copy = nullptr;
  int adj_blocks = 0;
  double *d_gout = nullptr;        // 6 model-grid arrays
  double *d_J = nullptr;
  int groups = 0;                  // shot groups of the last gradient (device memory)

  ~biot_batch();
  void drop_graph()
  {
    if (gexec) (void)hipGraphExecDestroy(gexec);
    gexec = nullptr;
    g_steps = 0;
  }
  void free_tables();
  void free_gradient();
};

namespace {

int bfail(biot_batch *b, int code, const char *fmt, ...)
{
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  b->err = buf;
  return code;
}

#define BHIP(b, call)                                                                      \
  do {                                                                                     \
    hipError_t e_ = (call);                                                                \
    if (e_ != hipSuccess)                                                                  \
      return bfail((b), BIOT_ERR_HIP, "%s failed: %s", #call, hipGetErrorString(e_));     \
  } while (0)

template <typename T>
void dfree(T *&p)
{
  if (p) (void)hipFree((void *)p);
  p = nullptr;
}

// shot-handle failures carry the handle's message
int hfail(biot_batch *b, biot_handle *h, int rc, const char *what)
{
  const char *m = h ? biot_last_error(h) : nullptr;
  return bfail(b, rc, "%s: %s", what, m && *m ? m : biot_status_string(rc));
}

}  // namespace

void biot_batch::free_tables()
{
  dfree(d_step);
  dfree(d_tiles);
  dfree(d_src);
  dfree(d_rec);
  built = false;
  drop_graph();
}

void biot_batch::free_gradient()
{
  if (stream) (void)hipStreamSynchronize(stream);
  for (auto &s : shots) {
    for (void *p : s.ckpt)
      if (p) (void)hipFree(p);
    s.ckpt.clear();
    for (auto &p : s.d_par) dfree(p);
    dfree(s.d_obs);
    if (s.h) s.h->free_adjoint();
  }
  dfree(d_adj);
  dfree(d_blocks);
  dfree(d_inj);
  dfree(d_copy);
  dfree(d_gout);
  dfree(d_J);
}

biot_batch::~biot_batch()
{
  free_gradient();
  free_tables();
  dfree(d_it);
  dfree(d_queue);
  for (auto &s : shots)
    if (s.h) biot_destroy(s.h);
  if (ev0) (void)hipEventDestroy(ev0);
  if (ev1) (void)hipEventDestroy(ev1);
  if (stream) (void)hipStreamDestroy(stream);
}

// message of the calling thread's last failed biot_batch_create (biot_batch_last_error(NULL))
static thread_local std::string t_batch_create_err;

extern "C" int biot_batch_create(const biot_batch_config *c, biot_batch **out)
{
  t_batch_create_err.clear();
  auto refuse = [](biot_batch *b, int rc, const char *why) {
    delete b;
    t_batch_create_err = why;
    return rc;
  };
  if (!c || !out) return refuse(nullptr, BIOT_ERR_ARG, "config and output pointers are required");
  *out = nullptr;
  if (c->nx < 8 || c->ny < 8 || c->nz < 8)
    return refuse(nullptr, BIOT_ERR_ARG, "the model grid must be at least 8 cells on every axis");
  if (c->order != 2 && c->order != 4) return refuse(nullptr, BIOT_ERR_ARG, "order must be 2 or 4");
  if (c->precision != BIOT_FP32 && c->precision != BIOT_FP16 && c->precision != BIOT_FP16T &&
      c->precision != BIOT_FP16C && c->precision != BIOT_FP16M && c->precision != BIOT_FP64)
    return refuse(nullptr, BIOT_ERR_ARG,
                  "batches run BIOT_FP32, BIOT_FP16, BIOT_FP16T, BIOT_FP16C, BIOT_FP16M or BIOT_FP64");
  if (!(c->dx > 0) || !(c->dt > 0) || c->npml < 0)
    return refuse(nullptr, BIOT_ERR_ARG, "dx and dt must be positive and npml >= 0");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return refuse(nullptr, BIOT_ERR_HIP, "no HIP device");
  if (c->device < 0 || c->device >= ndev) return refuse(nullptr, BIOT_ERR_ARG, "device index out of range");
  biot_batch *b = new biot_batch();
  b->cfg = *c;
  if (hipSetDevice(c->device) != hipSuccess || hipStreamCreate(&b->stream) != hipSuccess ||
      hipEventCreate(&b->ev0) != hipSuccess || hipEventCreate(&b->ev1) != hipSuccess ||
      hipMalloc((void **)&b->d_it, sizeof(int)) != hipSuccess ||
      hipMalloc((void **)&b->d_queue, 16 * sizeof(int)) != hipSuccess ||
      hipMemset(b->d_queue, 0, 16 * sizeof(int)) != hipSuccess)
    return refuse(b, BIOT_ERR_HIP, "device, stream, event or counter set-up failed");
  *out = b;
  return BIOT_OK;
}

extern "C" int biot_batch_destroy(biot_batch *b)
{
  if (!b) return BIOT_ERR_ARG;
  delete b;
  return BIOT_OK;
}

extern "C" const char *biot_batch_last_error(const biot_batch *b)
{
  return b ? b->err.c_str() : t_batch_create_err.c_str();
}

extern "C" int biot_batch_add_shot(biot_batch *b, const biot_shot *s, const biot_params *p,
                                   const float *wavelet, int32_t nt)
{
  if (!b) return BIOT_ERR_ARG;
  if (!s || !p || !wavelet || nt <= 0) return bfail(b, BIOT_ERR_ARG, "shot, parameters and wavelet required");
  const biot_batch_config &c = b->cfg;
  if (s->x0 < 0 || s->y0 < 0 || s->z0 < 0 || s->nx < 8 || s->ny < 8 || s->nz < 8 ||
      s->x0 + s->nx > c.nx || s->y0 + s->ny > c.ny || s->z0 + s->nz > c.nz)
    return bfail(b, BIOT_ERR_ARG, "window %dx%dx%d at (%d,%d,%d) outside the %dx%dx%d model", s->nx,
                 s->ny, s->nz, s->x0, s->y0, s->z0, c.nx, c.ny, c.nz);
  biot_grid g{};
  g.nx = s->nx;
  g.ny = s->ny;
  g.nz = s->nz;
  g.order = c.order;
  g.precision = c.precision;
  g.npml = c.npml;
  g.dx = c.dx;
  g.dt = c.dt;
  g.pml_v0 = c.pml_v0;
  g.pml_R = c.pml_R;
  g.Cv = c.Cv;
  g.Cp = c.Cp;
  g.device = c.device;
  biot_handle *h = nullptr;
  int rc = biot_create(&g, &h);
  if (rc) return bfail(b, rc, "shot %zu: window grid rejected: %s", b->shots.size(), biot_last_error(nullptr));
  if ((rc = biot_set_params(h, p)) || (rc = biot_set_source(h, s->src_cell, s->src_mask, wavelet, nt)) ||
      (rc = biot_set_receivers(h, s->rec_cells, s->nrec, s->rec_mask, nt))) {
    const int r2 = hfail(b, h, rc, "shot set-up");
    biot_destroy(h);
    return r2;
  }
  biot_batch::Shot sh;
  sh.h = h;
  sh.x0 = s->x0;
  sh.y0 = s->y0;
  sh.z0 = s->z0;
  b->free_gradient();
  b->free_tables();
  b->shots.push_back(sh);
  return BIOT_OK;
}

// New model-grid parameters: each shot's window is cut out and its coefficients rebuilt.
extern "C" int biot_batch_set_model(biot_batch *b, const biot_params *p)
{
  if (!b) return BIOT_ERR_ARG;
  if (!p || !p->lambda || !p->mu || !p->alpha || !p->M || !p->rho || !p->rhof || !p->C1 || !p->tphi)
    return bfail(b, BIOT_ERR_ARG, "every model parameter array is required");
  const double *src[8] = {p->lambda, p->mu, p->alpha, p->M, p->rho, p->rhof, p->C1, p->tphi};
  const int64_t msy = b->cfg.nx, msz = (int64_t)b->cfg.nx * b->cfg.ny;
  b->free_gradient();
  for (size_t i = 0; i < b->shots.size(); ++i) {
    auto &s = b->shots[i];
    biot_handle *h = s.h;
    const int wx = h->g.nx, wy = h->g.ny, wz = h->g.nz;
    std::vector<double> win[8];
    for (int q = 0; q < 8; ++q) {
      win[q].resize((size_t)h->n);
      for (int k = 0; k < wz; ++k)
        for (int j = 0; j < wy; ++j) {
          const double *row = src[q] + (s.x0 + (int64_t)(s.y0 + j) * msy + (int64_t)(s.z0 + k) * msz);
          std::memcpy(win[q].data() + ((int64_t)k * wy + j) * wx, row, (size_t)wx * sizeof(double));
        }
    }
    const biot_params wp{win[0].data(), win[1].data(), win[2].data(), win[3].data(),
                         win[4].data(), win[5].data(), win[6].data(), win[7].data()};
    int rc = biot_set_params(h, &wp);
    if (rc) return hfail(b, h, rc, "shot parameters");
  }
  b->built = false;
  b->drop_graph();
  return BIOT_OK;
}

extern "C" int biot_batch_shot_count(biot_batch *b, int32_t *n)
{
  if (!b || !n) return BIOT_ERR_ARG;
  *n = (int32_t)b->shots.size();
  return BIOT_OK;
}

namespace {

// device tables of the forward launches for the current shape, for shots [s0, s1) (all
// shots when s1 < 0); the tables hold the subset from their first entry
int build_tables(biot_batch *b, int s0 = 0, int s1 = -1)
{
  const bool all = s1 < 0;
  if (all && b->built) return BIOT_OK;
  if (all) s1 = (int)b->shots.size();
  const int S = s1 - s0;
  if (S <= 0) return bfail(b, BIOT_ERR_STATE, "no shots");
  const int Sall = (int)b->shots.size();
  std::vector<StepArgs> st(S);
  std::vector<int> tiles(S + 1, 0);
  std::vector<SourceArgs> src(S);
  std::vector<RecordArgs> rec(S);
  b->max_nrec = 0;
  b->max_ncomp = 0;
  const int h2 = b->cfg.order / 2;
  for (int i = 0; i < S; ++i) {
    biot_handle *h = b->shots[s0 + i].h;
    StepArgs &a = st[i];
    h->fill_args(a);
    a.zchunk = b->shape.zchunk;
    a.gx = (a.nx + b->shape.bx - 1) / b->shape.bx;
    a.gy = (a.ny + b->shape.by - 1) / b->shape.by;
    a.gz = (a.nz - 2 * h2 + a.zchunk - 1) / a.zchunk;
    a.ntiles = a.gx * a.gy * a.gz;
    tiles[i + 1] = tiles[i] + a.ntiles;
    h->source_args(src[i]);
    h->record_args(rec[i]);
    b->max_nrec = std::max(b->max_nrec, h->nrec);
    b->max_ncomp = std::max(b->max_ncomp, h->ncomp);
  }
  b->total_tiles = tiles[S];
  b->nactive = S;
  if (!b->d_step) {
    BHIP(b, hipMalloc((void **)&b->d_step, Sall * sizeof(StepArgs)));
    BHIP(b, hipMalloc((void **)&b->d_tiles, (Sall + 1) * sizeof(int)));
    BHIP(b, hipMalloc((void **)&b->d_src, Sall * sizeof(SourceArgs)));
    BHIP(b, hipMalloc((void **)&b->d_rec, Sall * sizeof(RecordArgs)));
  }
  BHIP(b, hipMemcpy(b->d_step, st.data(), S * sizeof(StepArgs), hipMemcpyHostToDevice));
  BHIP(b, hipMemcpy(b->d_tiles, tiles.data(), (S + 1) * sizeof(int), hipMemcpyHostToDevice));
  BHIP(b, hipMemcpy(b->d_src, src.data(), S * sizeof(SourceArgs), hipMemcpyHostToDevice));
  BHIP(b, hipMemcpy(b->d_rec, rec.data(), S * sizeof(RecordArgs), hipMemcpyHostToDevice));
  b->built = all;
  if (!all) b->drop_graph();  // a captured graph refers to the full tables
  return BIOT_OK;
}

BatchArgs batch_args(const biot_batch *b)
{
  // a grid smaller than the tile list takes its tiles from the work queue
  int *q = b->blocks < b->total_tiles ? b->d_queue : nullptr;
  return BatchArgs{b->d_step, b->d_tiles, b->nactive, b->total_tiles, q};
}

// one step of every shot: source, velocity, stress, (record), counter + 1
int enqueue_step(biot_batch *b, bool record)
{
  const int S = b->nactive;
  const int prec = b->cfg.precision, order = b->cfg.order;
  const BatchArgs ba = batch_args(b);
  BHIP(b, launch_source_batch(prec, b->d_src, S, b->d_it, b->stream));
  BHIP(b, launch_step_batch(prec, order, true, ba, b->shape, b->blocks, b->stream));
  BHIP(b, launch_step_batch(prec, order, false, ba, b->shape, b->blocks, b->stream));
  if (record) BHIP(b, launch_record_batch(prec, b->d_rec, S, b->max_nrec, b->max_ncomp, b->d_it, b->stream));
  BHIP(b, launch_advance(b->d_it, b->stream));
  return BIOT_OK;
}

int set_counter(biot_batch *b, int it)
{
  BHIP(b, launch_set_counter(b->d_it, it, b->stream));
  return BIOT_OK;
}

int ensure_graph(biot_batch *b, int steps, bool record)
{
  if (b->gexec && b->g_steps == steps && b->g_record == record) return BIOT_OK;
  b->drop_graph();
  hipGraph_t graph = nullptr;
  BHIP(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
  int rc = BIOT_OK;
  for (int s = 0; s < steps && !rc; ++s) rc = enqueue_step(b, record);
  hipGraph_t g2 = nullptr;
  const hipError_t e = hipStreamEndCapture(b->stream, &g2);
  graph = g2;
  if (rc) {
    if (graph) (void)hipGraphDestroy(graph);
    return rc;
  }
  if (e != hipSuccess) return bfail(b, BIOT_ERR_HIP, "graph capture: %s", hipGetErrorString(e));
  const hipError_t e2 = hipGraphInstantiate(&b->gexec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (e2 != hipSuccess) {
    b->gexec = nullptr;
    return bfail(b, BIOT_ERR_HIP, "graph instantiate: %s", hipGetErrorString(e2));
  }
  b->g_steps = steps;
  b->g_record = record;
  return BIOT_OK;
}

// nt steps from the counter's current value: whole graphs, then single steps
int run_steps(biot_batch *b, int nt, int graph_steps, bool record)
{
  int done = 0;
  if (graph_steps > 0 && nt >= graph_steps) {
    int rc = ensure_graph(b, graph_steps, record);
    if (rc) return rc;
    for (; done + graph_steps <= nt; done += graph_steps) BHIP(b, hipGraphLaunch(b->gexec, b->stream));
  }
  for (; done < nt; ++done) {
    int rc = enqueue_step(b, record);
    if (rc) return rc;
  }
  return BIOT_OK;
}

int reset_all(biot_batch *b)
{
  for (auto &s : b->shots) {
    int rc = biot_reset(s.h);
    if (rc) return hfail(b, s.h, rc, "reset");
  }
  return BIOT_OK;
}

}  // namespace

extern "C" int biot_batch_set_launch(biot_batch *b, int32_t bx, int32_t by, int32_t zchunk, int32_t blocks)
{
  if (!b) return BIOT_ERR_ARG;
  if (bx < 64 || bx % 64 || by < 1 || bx * by > 256 || zchunk < 1 || blocks < 8 || blocks % 8)
    return bfail(b, BIOT_ERR_ARG, "launch %dx%d, zchunk %d, %d blocks not allowed", bx, by, zchunk, blocks);
  b->shape = LaunchShape{bx, by, zchunk};
  b->blocks = blocks;
  b->built = false;
  b->drop_graph();
  return BIOT_OK;
}

extern "C" int biot_batch_get_launch(biot_batch *b, int32_t *bx, int32_t *by, int32_t *zchunk, int32_t *blocks)
{
  if (!b || !bx || !by || !zchunk || !blocks) return BIOT_ERR_ARG;
  *bx = b->shape.bx;
  *by = b->shape.by;
  *zchunk = b->shape.zchunk;
  *blocks = b->blocks;
  return BIOT_OK;
}

extern "C" int biot_batch_autotune(biot_batch *b, int32_t trial_steps)
{
  if (!b) return BIOT_ERR_ARG;
  if (trial_steps < 1 || trial_steps > 1000) return bfail(b, BIOT_ERR_ARG, "trial_steps must be 1..1000");
  if (b->shots.empty()) return bfail(b, BIOT_ERR_STATE, "no shots");
  static const int shapes[][2] = {{64, 4}, {64, 2}, {64, 1}, {128, 2}, {128, 1}, {256, 1}};
  static const int chunks[] = {2, 4, 8, 16, 32};
  static const int grids[] = {1024, 2048, 4096, 1 << 30};
  const LaunchShape keep = b->shape;
  const int keep_blocks = b->blocks;
  LaunchShape best = keep;
  int best_blocks = keep_blocks;
  double best_ms = 1e30;
  int rc = reset_all(b);
  for (auto &sh : shapes)
    for (int zc : chunks)
      for (int gblocks : grids) {
        if (rc) break;
        b->shape = LaunchShape{sh[0], sh[1], zc};
        b->built = false;
        b->drop_graph();
        if ((rc = build_tables(b))) break;
        const int gb = std::min(gblocks, ((b->total_tiles + 7) / 8) * 8);
        if (gblocks == (1 << 30) || gb == gblocks) b->blocks = gb;
        else continue;  // a grid larger than the tile list is the "one block per tile" case
        if ((rc = set_counter(b, 0))) break;
        if ((rc = enqueue_step(b, false))) break;  // warm-up
        float ms = 0;
        if (hipEventRecord(b->ev0, b->stream) != hipSuccess) rc = BIOT_ERR_HIP;
        for (int s = 0; s < trial_steps && !rc; ++s) rc = enqueue_step(b, false);
        if (!rc && (hipEventRecord(b->ev1, b->stream) != hipSuccess || hipEventSynchronize(b->ev1) != hipSuccess ||
                    hipEventElapsedTime(&ms, b->ev0, b->ev1) != hipSuccess))
          rc = BIOT_ERR_HIP;
        if (!rc && ms < best_ms) {
          best_ms = ms;
          best = b->shape;
          best_blocks = b->blocks;
        }
      }
  b->shape = rc ? keep : best;
  b->blocks = rc ? keep_blocks : best_blocks;
  b->built = false;
  b->drop_graph();
  if (rc) return rc == BIOT_ERR_HIP ? bfail(b, rc, "autotune timing failed") : rc;
  return reset_all(b);
}

extern "C" int biot_batch_forward(biot_batch *b, int32_t nt, int32_t graph_steps)
{
  if (!b) return BIOT_ERR_ARG;
  if (nt <= 0 || graph_steps < 0) return bfail(b, BIOT_ERR_ARG, "nt must be positive, graph_steps >= 0");
  int rc = build_tables(b);
  if (!rc) rc = reset_all(b);
  if (!rc) rc = set_counter(b, 0);
  if (rc) return rc;
  BHIP(b, hipEventRecord(b->ev0, b->stream));
  rc = run_steps(b, nt, graph_steps, true);
  if (rc) return rc;
  BHIP(b, hipEventRecord(b->ev1, b->stream));
  BHIP(b, hipEventSynchronize(b->ev1));
  float ms = 0;
  BHIP(b, hipEventElapsedTime(&ms, b->ev0, b->ev1));
  b->last_ms = ms;
  for (auto &s : b->shots) s.h->it = nt;
  for (size_t i = 0; i < b->shots.size(); ++i)
    if ((rc = b->shots[i].h->check_nonfinite())) {
      char what[32];
      snprintf(what, sizeof what, "shot %d", (int)i);
      return hfail(b, b->shots[i].h, rc, what);
    }
  return BIOT_OK;
}

extern "C" int biot_batch_read_traces(biot_batch *b, int32_t shot, float *out, int64_t count)
{
  if (!b || !out) return BIOT_ERR_ARG;
  if (shot < 0 || shot >= (int)b->shots.size()) return bfail(b, BIOT_ERR_ARG, "no shot %d", shot);
  biot_handle *h = b->shots[shot].h;
  int rc = biot_read_traces(h, out, count);
  return rc ? hfail(b, h, rc, "read traces") : BIOT_OK;
}

extern "C" int biot_batch_shot_handle(biot_batch *b, int32_t shot, biot_handle **h)
{
  if (!b || !h) return BIOT_ERR_ARG;
  if (shot < 0 || shot >= (int)b->shots.size()) return bfail(b, BIOT_ERR_ARG, "no shot %d", shot);
  *h = b->shots[shot].h;
  return BIOT_OK;
}

namespace {

// device bytes the gradient needs for one shot
size_t shot_gradient_bytes(const biot_handle *h, int nseg, int every)
{
  const int64_t m = (int64_t)h->ncomp * h->nt_max * h->nrec;
  return h->state_bytes() * (size_t)(nseg + every + 1) + h->lam_bytes() +
         (size_t)NKC * h->n * (sizeof(double) + (h->prec == BIOT_FP64 ? 0 : sizeof(float))) +
         (h->prec == BIOT_FP16M ? (size_t)NKC * h->n * sizeof(_Float16) : 0) +  // adjoint coefficients
         (size_t)8 * h->n * sizeof(double) +
         (size_t)2 * std::max<int64_t>(m, 1) * sizeof(float);
}

// the gradient of shots [s0, s1): checkpointed forward, misfit, recompute + adjoint, chain
int gradient_group(biot_batch *b, int s0, int s1, int nt, int every, const float *const observed[])
{
  const int S = s1 - s0;
  const int prec = b->cfg.precision, order = b->cfg.order;
  const int nseg = (nt + every - 1) / every;
  const int64_t N = (int64_t)b->cfg.nx * b->cfg.ny * b->cfg.nz;
  int rc = build_tables(b, s0, s1);
  if (rc) return rc;
  for (int i = s0; i < s1; ++i) {
    auto &s = b->shots[i];
    biot_handle *h = s.h;
    if ((rc = h->alloc_adjoint(every))) return hfail(b, h, rc, "adjoint buffers");
    if ((rc = h->build_adjoint_coefs())) return hfail(b, h, rc, "adjoint coefficients");
    for (int k = 0; k < nseg; ++k) {
      void *p = nullptr;
      if (hipMalloc(&p, h->state_bytes()) != hipSuccess)
        return bfail(b, BIOT_ERR_NOMEM, "checkpoints of shot %d do not fit", i);
      s.ckpt.push_back(p);
    }
    for (int q = 0; q < 8; ++q) {
      BHIP(b, hipMalloc((void **)&s.d_par[q], (size_t)h->n * sizeof(double)));
      BHIP(b, hipMemcpy(s.d_par[q], h->par[q].data(), (size_t)h->n * sizeof(double), hipMemcpyHostToDevice));
    }
    const int64_t m = (int64_t)h->ncomp * h->nt_max * h->nrec;
    BHIP(b, hipMalloc((void **)&s.d_obs, (size_t)std::max<int64_t>(m, 1) * sizeof(float)));
    BHIP(b, hipMemcpy(s.d_obs, observed[i], (size_t)m * sizeof(float), hipMemcpyHostToDevice));
  }
  // adjoint / inject tables
  {
    std::vector<AdjArgs> adj(S);
    std::vector<InjectArgs> inj(S);
    std::vector<int> bs(S + 1, 0);
    for (int i = 0; i < S; ++i) {
      b->shots[s0 + i].h->adj_args(adj[i]);
      b->shots[s0 + i].h->inject_args(inj[i]);
      bs[i + 1] = bs[i] + (int)((adj[i].nupd + 255) / 256);
    }
    b->adj_blocks = bs[S];
    BHIP(b, hipMalloc((void **)&b->d_adj, S * sizeof(AdjArgs)));
    BHIP(b, hipMalloc((void **)&b->d_blocks, (S + 1) * sizeof(int)));
    BHIP(b, hipMalloc((void **)&b->d_inj, S * sizeof(InjectArgs)));
    BHIP(b, hipMemcpy(b->d_adj, adj.data(), S * sizeof(AdjArgs), hipMemcpyHostToDevice));
    BHIP(b, hipMemcpy(b->d_blocks, bs.data(), (S + 1) * sizeof(int), hipMemcpyHostToDevice));
    BHIP(b, hipMemcpy(b->d_inj, inj.data(), S * sizeof(InjectArgs), hipMemcpyHostToDevice));
  }
  // copy tables: [save ckpt k] nseg, [load ckpt k] nseg, [save slot q] every + 1; 5 pieces per shot
  // (fields, the three carry blocks, the FP16T / FP16C residuals: 0 bytes on the other policies)
  constexpr int NP = 5;
  const int per = NP * S;
  auto pieces = [&](biot_handle *h, char *state, bool to_state, CopyEntry *e) {
    char *live[NP] = {(char *)h->d_fields, (char *)h->d_carry[0], (char *)h->d_carry[1], (char *)h->d_carry[2],
                      (char *)h->d_lo};
    const size_t bytes[NP] = {(size_t)NFIELD * h->n * h->esz, (size_t)NCARRY * h->slab[0] * h->csz,
                              (size_t)NCARRY * h->slab[1] * h->csz, (size_t)NCARRY * h->slab[2] * h->csz,
                              h->d_lo ? h->lo_bytes() : 0};
    for (int p = 0; p < NP; ++p) {
      char *st = state + (p < 4 ? h->state_off(p) : h->state_lo_off());
      e[p].src = to_state ? (const void *)live[p] : (const void *)st;
      e[p].dst = to_state ? (void *)st : (void *)live[p];
      e[p].bytes = bytes[p];
    }
  };
  {
    std::vector<CopyEntry> tab((size_t)(2 * nseg + every + 1) * per);
    for (int i = 0; i < S; ++i) {
      biot_handle *h = b->shots[s0 + i].h;
      for (int k = 0; k < nseg; ++k) {
        pieces(h, (char *)b->shots[s0 + i].ckpt[k], true, &tab[(size_t)k * per + NP * i]);
        pieces(h, (char *)b->shots[s0 + i].ckpt[k], false, &tab[(size_t)(nseg + k) * per + NP * i]);
      }
      for (int q = 0; q <= every; ++q)
        pieces(h, (char *)h->d_slots + (size_t)q * h->state_bytes(), true, &tab[(size_t)(2 * nseg + q) * per + NP * i]);
    }
    BHIP(b, hipMalloc((void **)&b->d_copy, tab.size() * sizeof(CopyEntry)));
    BHIP(b, hipMemcpy(b->d_copy, tab.data(), tab.size() * sizeof(CopyEntry), hipMemcpyHostToDevice));
  }
  auto save_ckpt = [&](int k) { return launch_copy_batch(b->d_copy + (size_t)k * per, per, b->stream); };
  auto load_ckpt = [&](int k) { return launch_copy_batch(b->d_copy + (size_t)(nseg + k) * per, per, b->stream); };
  auto save_slot = [&](int q) { return launch_copy_batch(b->d_copy + (size_t)(2 * nseg + q) * per, per, b->stream); };
  for (int i = s0; i < s1; ++i) {
    int r = biot_reset(b->shots[i].h);
    if (r) return hfail(b, b->shots[i].h, r, "reset");
  }
  if ((rc = set_counter(b, 0))) return rc;

  // ---- forward with checkpoints
  for (int k = 0; k < nseg; ++k) {
    BHIP(b, save_ckpt(k));
    if ((rc = run_steps(b, std::min(every, nt - k * every), 0, true))) return rc;
  }
  // ---- FP64 misfit per shot; the residual is the adjoint source
  for (int i = s0; i < s1; ++i) {
    biot_handle *h = b->shots[i].h;
    const int64_t m = (int64_t)h->ncomp * h->nt_max * h->nrec;
    const bool dbl = h->d_traces64 != nullptr;
    BHIP(b, launch_misfit(dbl ? (const void *)h->d_traces64 : h->d_traces, dbl, b->shots[i].d_obs,
                          h->d_adjsrc, m, b->d_J, b->stream));
    BHIP(b, hipMemsetAsync(h->d_lam, 0, h->lam_bytes(), b->stream));
    BHIP(b, hipMemsetAsync(h->d_grad, 0, (size_t)NKC * h->n * sizeof(double), b->stream));
    if (h->d_gpart) BHIP(b, hipMemsetAsync(h->d_gpart, 0, (size_t)NKC * h->n * sizeof(float), b->stream));
  }
  // ---- segment recompute and adjoint sweep, all shots of the group per launch
  const AdjBatch ab{b->d_adj, b->d_blocks, S, b->adj_blocks};
  for (int k = nseg - 1; k >= 0; --k) {
    const int n0 = k * every, K = std::min(every, nt - n0);
    BHIP(b, load_ckpt(k));
    BHIP(b, launch_set_counter(b->d_it, n0, b->stream));
    const BatchArgs ba = batch_args(b);
    for (int q = 0; q < K; ++q) {
      BHIP(b, launch_source_batch(prec, b->d_src, S, b->d_it, b->stream));
      BHIP(b, save_slot(q));
      BHIP(b, launch_step_batch(prec, order, true, ba, b->shape, b->blocks, b->stream));
      BHIP(b, launch_step_batch(prec, order, false, ba, b->shape, b->blocks, b->stream));
      BHIP(b, launch_advance(b->d_it, b->stream));
    }
    BHIP(b, save_slot(K));
    for (int it = n0 + K - 1; it >= n0; --it) {
      BHIP(b, launch_adjoint_inject_batch(prec, b->d_inj, S, b->max_nrec, b->max_ncomp, it, b->stream));
      BHIP(b, launch_adjoint_step_batch(prec, order, ab, it - n0, b->stream));
    }
    BHIP(b, launch_adjoint_flush_batch(prec, order, ab, b->stream));
  }
  // ---- chain rule per shot into the model-grid gradient
  for (int i = s0; i < s1; ++i) {
    auto &s = b->shots[i];
    biot_handle *h = s.h;
    ChainArgs c{};
    for (int k = 0; k < NKC; ++k) {
      c.grad[k] = h->d_grad + (size_t)k * h->n;
      c.kfac[k] = h->kfac[k];
    }
    for (int q = 0; q < 8; ++q) c.par[q] = s.d_par[q];
    for (int q = 0; q < 6; ++q) c.out[q] = b->d_gout + (size_t)q * N;
    c.wx = h->g.nx;
    c.wy = h->g.ny;
    c.wz = h->g.nz;
    c.x0 = s.x0;
    c.y0 = s.y0;
    c.z0 = s.z0;
    c.msy = b->cfg.nx;
    c.msz = (int64_t)b->cfg.nx * b->cfg.ny;
    c.dt = b->cfg.dt;
    c.Cv = b->cfg.Cv;
    c.Cp = b->cfg.Cp;
    c.Hinv = (double)(1.0f / b->cfg.dx);
    BHIP(b, launch_chain(c, b->stream));
  }
  for (int i = s0; i < s1; ++i) b->shots[i].h->it = nt;
  return BIOT_OK;
}

// releases the per-shot gradient buffers of shots [s0, s1) and the group tables
void free_group(biot_batch *b, int s0, int s1)
{
  (void)hipStreamSynchronize(b->stream);
  for (int i = s0; i < s1; ++i) {
    auto &s = b->shots[i];
    for (void *p : s.ckpt)
      if (p) (void)hipFree(p);
    s.ckpt.clear();
    for (auto &p : s.d_par) dfree(p);
    dfree(s.d_obs);
    s.h->free_adjoint();
  }
  dfree(b->d_adj);
  dfree(b->d_blocks);
  dfree(b->d_inj);
  dfree(b->d_copy);
}

}  // namespace

extern "C" int biot_batch_gradient(biot_batch *b, int32_t nt, int32_t every, const float *const observed[],
                                   double *J, double *const grad[6])
{
  if (!b) return BIOT_ERR_ARG;
  if (!observed || !J || !grad) return bfail(b, BIOT_ERR_ARG, "observed, J and grad are required");
  for (int q = 0; q < 6; ++q)
    if (!grad[q]) return bfail(b, BIOT_ERR_ARG, "grad[%d] is null", q);
  const int S = (int)b->shots.size();
  if (S == 0) return bfail(b, BIOT_ERR_STATE, "no shots");
  if (nt <= 0 || every <= 0) return bfail(b, BIOT_ERR_ARG, "nt and every must be positive");
  for (int i = 0; i < S; ++i) {
    if (!observed[i]) return bfail(b, BIOT_ERR_ARG, "observed[%d] is null", i);
    if (b->shots[i].h->nt_max < nt) return bfail(b, BIOT_ERR_ARG, "shot %d records fewer than nt steps", i);
    if (b->shots[i].h->src_mask & 0x3Fu)
      return bfail(b, BIOT_ERR_ARG, "shot %d: the adjoint path supports sources on stresses and pressure only", i);
  }
  const int nseg = (nt + every - 1) / every;
  const int64_t N = (int64_t)b->cfg.nx * b->cfg.ny * b->cfg.nz;
  b->free_gradient();
  BHIP(b, hipMalloc((void **)&b->d_gout, (size_t)6 * N * sizeof(double)));
  BHIP(b, hipMalloc((void **)&b->d_J, sizeof(double)));
  // shots go in groups that fit the free device memory (one group when they all fit)
  size_t free_b = 0, total_b = 0;
  BHIP(b, hipMemGetInfo(&free_b, &total_b));
  const size_t budget = (size_t)((double)free_b * 0.9);
  std::vector<std::pair<int, int>> groups;
  for (int i = 0; i < S;) {
    size_t used = 0;
    int j = i;
    while (j < S) {
      const size_t need = shot_gradient_bytes(b->shots[j].h, nseg, every);
      if (j > i && used + need > budget) break;
      used += need;
      ++j;
    }
    groups.emplace_back(i, j);
    i = j;
  }
  BHIP(b, hipMemsetAsync(b->d_gout, 0, (size_t)6 * N * sizeof(double), b->stream));
  BHIP(b, hipMemsetAsync(b->d_J, 0, sizeof(double), b->stream));
  BHIP(b, hipEventRecord(b->ev0, b->stream));
  int rc = BIOT_OK;
  for (auto &g : groups) {
    rc = gradient_group(b, g.first, g.second, nt, every, observed);
    free_group(b, g.first, g.second);
    if (rc) break;
  }
  b->built = false;
  if (rc) return rc;
  BHIP(b, hipEventRecord(b->ev1, b->stream));
  BHIP(b, hipEventSynchronize(b->ev1));
  float ms = 0;
  BHIP(b, hipEventElapsedTime(&ms, b->ev0, b->ev1));
  b->last_ms = ms;
  b->groups = (int)groups.size();
  BHIP(b, hipMemcpy(J, b->d_J, sizeof(double), hipMemcpyDeviceToHost));
  for (int q = 0; q < 6; ++q)
    BHIP(b, hipMemcpy(grad[q], b->d_gout + (size_t)q * N, (size_t)N * sizeof(double), hipMemcpyDeviceToHost));
  return BIOT_OK;
}

extern "C" int biot_batch_gradient_groups(biot_batch *b, int32_t *groups)
{
  if (!b || !groups) return BIOT_ERR_ARG;
  *groups = b->groups;
  return BIOT_OK;
}

extern "C" int biot_batch_last_ms(biot_batch *b, double *ms)
{
  if (!b || !ms) return BIOT_ERR_ARG;
  *ms = b->last_ms;
  return BIOT_OK;
}

extern "C" int biot_batch_release(biot_batch *b)
{
  if (!b) return BIOT_ERR_ARG;
  b->free_gradient();
  return BIOT_OK;
}
