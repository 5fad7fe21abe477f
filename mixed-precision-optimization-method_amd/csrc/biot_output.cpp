// Trace and snapshot output: the reference's text writers (outputfunc.cpp:35-50, 97-151)
// as host-only calls, a raw binary writer, and a time loop that streams the recorded traces
// to host memory on a side stream while the stencil launches continue (the reference copies
// the whole volume to the host every step, kernel.cu:2197-2211, and writes the files after
// the loop, kernel.cu:2557-2571).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/biot.h"
#include "biot_handle.hpp"

using namespace biot;

namespace {

// one row: "%e " per value, the row's newline, then the empty line the reference writes
bool write_row(FILE *fp, const float *v, int64_t n, int64_t stride)
{
  for (int64_t k = 0; k < n; ++k)
    if (std::fprintf(fp, "%e ", (double)v[k * stride]) < 0) return false;
  return std::fputs("\n\n", fp) >= 0;
}

}  // namespace

extern "C" int biot_write_text_2d(const char *path, const float *data, int32_t nr, int32_t nc)
{
  if (!path || (!data && nr > 0 && nc > 0) || nr < 0 || nc < 0) return BIOT_ERR_ARG;
  FILE *fp = std::fopen(path, "w");
  if (!fp) return BIOT_ERR_ARG;
  bool ok = true;
  for (int32_t i = 0; i < nr && ok; ++i) {
    if (nc == 0) ok = std::fputs("\n", fp) >= 0;
    else ok = write_row(fp, data + (int64_t)i * nc, nc, 1);
  }
  ok = (std::fclose(fp) == 0) && ok;
  return ok ? BIOT_OK : BIOT_ERR_ARG;
}

extern "C" int biot_write_text_xz(const char *path, const float *vol, int32_t nz, int32_t ny, int32_t nx)
{
  if (!path || !vol || nz <= 0 || ny <= 0 || nx <= 0) return BIOT_ERR_ARG;
  FILE *fp = std::fopen(path, "w");
  if (!fp) return BIOT_ERR_ARG;
  const int32_t j = ny / 2;  // the reference's y slice
  bool ok = true;
  for (int32_t k = 0; k < nz && ok; ++k)
    ok = write_row(fp, vol + ((int64_t)k * ny + j) * nx, nx, 1);
  ok = (std::fclose(fp) == 0) && ok;
  return ok ? BIOT_OK : BIOT_ERR_ARG;
}

extern "C" int biot_write_binary(const char *path, const float *data, int64_t count)
{
  if (!path || count < 0 || (!data && count > 0)) return BIOT_ERR_ARG;
  FILE *fp = std::fopen(path, "wb");
  if (!fp) return BIOT_ERR_ARG;
  const size_t w = count ? std::fwrite(data, sizeof(float), (size_t)count, fp) : 0;
  const bool ok = (std::fclose(fp) == 0) && w == (size_t)count;
  return ok ? BIOT_OK : BIOT_ERR_ARG;
}

// nsteps from the handle's current step; the recorded traces (if `traces`) are copied to the
// host chunk steps at a time and xz snapshots are taken after the listed steps, all on a side
// stream while the compute stream continues.  Trace chunks: the compute stream records an
// event after each chunk, the side stream waits for it and copies the chunk's rows of every
// component into one of two pinned staging buffers, and the host moves the previous chunk from
// staging into `traces` meanwhile.  Snapshots: after step it0 + snap_steps[s] the compute
// stream writes the slice (physical units, float) into a device buffer of its own, and the side
// stream copies it into pinned memory; the host copies it out at the end.
extern "C" int biot_run_stream(biot_handle *h, int32_t nsteps, int32_t chunk, float *traces, int32_t field,
                               int32_t j, const int32_t *snap_steps, int32_t nsnap, float *snaps)
{
  if (!h) return BIOT_ERR_ARG;
  if (nsteps < 0 || chunk <= 0 || nsnap < 0 || (nsnap > 0 && (!snap_steps || !snaps))) {
    h->err = "nsteps >= 0, chunk > 0, and snapshot steps and output for nsnap > 0 are required";
    return BIOT_ERR_ARG;
  }
  if (!h->have_coef) {
    h->err = "biot_run_stream before set_medium";
    return BIOT_ERR_STATE;
  }
  if (traces && (!h->d_traces || h->ncomp == 0)) {
    h->err = "no receivers";
    return BIOT_ERR_STATE;
  }
  if (nsnap > 0) {
    if (field < 0 || field >= NFIELD || j < 0 || j >= h->g.ny) {
      h->err = "snapshot field or y index out of range";
      return BIOT_ERR_ARG;
    }
    for (int s = 0; s < nsnap; ++s)
      if (snap_steps[s] < 0 || snap_steps[s] >= nsteps || (s > 0 && snap_steps[s] <= snap_steps[s - 1])) {
        h->err = "snapshot steps must ascend within [0, nsteps)";
        return BIOT_ERR_ARG;
      }
  }
  const int it0 = h->it;
  const int last = std::min(it0 + nsteps, h->nt_max);  // steps beyond nt_max record nothing
  chunk = std::min(chunk, std::max(nsteps, 1));       // the staging buffers hold one chunk
  const size_t row = (size_t)h->nrec * sizeof(float);
  const size_t stage_bytes = traces ? (size_t)h->ncomp * chunk * row : 0;
  const size_t slice = (size_t)h->g.nx * h->g.nz;
  hipStream_t side = nullptr;
  float *stage[2] = {nullptr, nullptr};
  float *d_snap = nullptr, *h_snap = nullptr;
  hipEvent_t done_compute[2] = {nullptr, nullptr}, done_copy[2] = {nullptr, nullptr};
  std::vector<hipEvent_t> snap_ev((size_t)nsnap, nullptr);
  int rc = BIOT_OK;
  auto cleanup = [&]() {
    if (side) (void)hipStreamSynchronize(side);
    for (hipEvent_t e : snap_ev)
      if (e) (void)hipEventDestroy(e);
    for (int b = 0; b < 2; ++b) {
      if (stage[b]) (void)hipHostFree(stage[b]);
      if (done_compute[b]) (void)hipEventDestroy(done_compute[b]);
      if (done_copy[b]) (void)hipEventDestroy(done_copy[b]);
    }
    if (h_snap) (void)hipHostFree(h_snap);
    if (d_snap) (void)hipFree(d_snap);
    if (side) (void)hipStreamDestroy(side);
  };
  if (hipStreamCreate(&side) != hipSuccess) {
    h->err = "side stream";
    return BIOT_ERR_HIP;
  }
  for (int b = 0; b < 2 && !rc; ++b)
    if ((traces && hipHostMalloc((void **)&stage[b], std::max<size_t>(stage_bytes, 4)) != hipSuccess) ||
        hipEventCreateWithFlags(&done_compute[b], hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&done_copy[b], hipEventDisableTiming) != hipSuccess)
      rc = BIOT_ERR_NOMEM;
  if (!rc && nsnap > 0 &&
      (hipMalloc((void **)&d_snap, (size_t)nsnap * slice * sizeof(float)) != hipSuccess ||
       hipHostMalloc((void **)&h_snap, (size_t)nsnap * slice * sizeof(float)) != hipSuccess))
    rc = BIOT_ERR_NOMEM;
  for (int s = 0; s < nsnap && !rc; ++s)
    if (hipEventCreateWithFlags(&snap_ev[s], hipEventDisableTiming) != hipSuccess) rc = BIOT_ERR_NOMEM;
  if (rc) {
    cleanup();
    h->err = "staging buffers";
    return rc;
  }
  // host side of trace chunk [a, b) in buffer s: wait for its copy, move rows into `traces`
  auto drain = [&](int s, int a, int b) -> bool {
    if (hipEventSynchronize(done_copy[s]) != hipSuccess) return false;
    for (int c = 0; c < h->ncomp; ++c)
      std::memcpy(traces + ((size_t)c * h->nt_max + a) * h->nrec, (char *)stage[s] + (size_t)c * chunk * row,
                  (size_t)(b - a) * row);
    return true;
  };
  int pend_s = -1, pend_a = 0, pend_b = 0;
  if (hipEventRecord(h->ev0, h->stream) != hipSuccess) rc = BIOT_ERR_HIP;
  int k = 0, next_snap = 0;
  for (int a = it0; a < it0 + nsteps && !rc;) {
    // a chunk ends after `chunk` steps, at the end, or right after the next snapshot step
    int b = std::min(a + chunk, it0 + nsteps);
    if (next_snap < nsnap) b = std::min(b, it0 + snap_steps[next_snap] + 1);
    rc = h->enqueue_steps(b - a, nullptr);
    if (rc) break;
    if (next_snap < nsnap && b == it0 + snap_steps[next_snap] + 1) {
      float *ds = d_snap + (size_t)next_snap * slice;
      if (launch_slice_xz(h->prec, h->field_ptr(field), h->lo_ptr(field), ((comp_fields(h->prec) >> field) & 1u) != 0,
                          (double)h->fscale[field], j, h->g.nx, h->g.ny, h->g.nz, ds, h->stream) != hipSuccess) {
        rc = BIOT_ERR_HIP;
        break;
      }
      // the side stream copies the slice once the compute stream has written it
      hipEvent_t ev = snap_ev[next_snap];
      if (hipEventRecord(ev, h->stream) != hipSuccess || hipStreamWaitEvent(side, ev, 0) != hipSuccess ||
          hipMemcpyAsync(h_snap + (size_t)next_snap * slice, ds, slice * sizeof(float), hipMemcpyDeviceToHost, side) !=
              hipSuccess)
        rc = BIOT_ERR_HIP;
      if (rc) break;
      ++next_snap;
    }
    const int ra = std::min(a, last), rb = std::min(b, last);  // recorded rows of this chunk
    if (traces && rb > ra) {
      const int s = k & 1;
      ++k;
      // the staging buffer s was last drained two chunks ago (drained below before reuse)
      if (hipEventRecord(done_compute[s], h->stream) != hipSuccess ||
          hipStreamWaitEvent(side, done_compute[s], 0) != hipSuccess) {
        rc = BIOT_ERR_HIP;
        break;
      }
      for (int c = 0; c < h->ncomp && !rc; ++c)
        if (hipMemcpyAsync((char *)stage[s] + (size_t)c * chunk * row,
                           (const char *)h->d_traces + ((size_t)c * h->nt_max + ra) * row,
                           (size_t)(rb - ra) * row, hipMemcpyDeviceToHost, side) != hipSuccess)
          rc = BIOT_ERR_HIP;
      if (!rc && hipEventRecord(done_copy[s], side) != hipSuccess) rc = BIOT_ERR_HIP;
      if (rc) break;
      // drain the previous chunk (other buffer) while this one computes and copies
      if (pend_s >= 0 && !drain(pend_s, pend_a, pend_b)) {
        rc = BIOT_ERR_HIP;
        break;
      }
      pend_s = s, pend_a = ra, pend_b = rb;
    }
    a = b;
  }
  if (!rc && pend_s >= 0 && !drain(pend_s, pend_a, pend_b)) rc = BIOT_ERR_HIP;
  if (!rc && (hipEventRecord(h->ev1, h->stream) != hipSuccess || hipEventSynchronize(h->ev1) != hipSuccess))
    rc = BIOT_ERR_HIP;
  if (!rc) {
    float ms = 0;
    if (hipEventElapsedTime(&ms, h->ev0, h->ev1) == hipSuccess) h->last_ms = ms;
  }
  if (!rc && nsnap > 0) {
    if (hipStreamSynchronize(side) != hipSuccess) rc = BIOT_ERR_HIP;
    else std::memcpy(snaps, h_snap, (size_t)nsnap * slice * sizeof(float));
  }
  if (!rc) rc = h->check_nonfinite();
  cleanup();
  if (rc == BIOT_ERR_HIP && h->err.empty()) h->err = "streamed copy failed";
  return rc;
}

extern "C" int biot_run_stream_traces(biot_handle *h, int32_t nsteps, int32_t chunk, float *out)
{
  if (!h) return BIOT_ERR_ARG;
  if (!out) {
    h->err = "an output array is required";
    return BIOT_ERR_ARG;
  }
  return biot_run_stream(h, nsteps, chunk, out, 0, 0, nullptr, 0, nullptr);
}
