// Restart files of the reference (write_restart / read_restart,
// src/CFD_2nd/m_IOAndVisu.f90:352-438, 443-558; src/CFD_4th/m_IOAndVisu.f90:
// 384-489): the interior points of ux, uy, uz, pressure (and optionally the
// three history fields), each a real(8) array in Fortran order, back to back.
//
// The interior rows of one z plane are contiguous in the device layout
// (rows jj = 1..ny of x-contiguous lines), so a run of planes moves with one
// strided 2-D copy.  Two pinned buffers alternate: the copy of the next run
// proceeds while the current one is written to (or read from) the file.
#include "c3d_internal.h"

#include <sys/stat.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>

namespace c3d {

namespace {

constexpr size_t kRunBytes = 64u << 20;  // pinned bytes per buffer

struct Pinned {
  explicit Pinned(hipStream_t s) : stream(s) {}
  hipStream_t stream;
  void* p[2] = {nullptr, nullptr};
  double* conv = nullptr;  // FP32 contexts: conversion buffer
  hipEvent_t ev[2] = {nullptr, nullptr};
  ~Pinned() {
    // no copy may still target the buffers when they are released
    hipStreamSynchronize(stream);
    for (int b = 0; b < 2; ++b) {
      if (p[b]) hipHostFree(p[b]);
      if (ev[b]) hipEventDestroy(ev[b]);
    }
    delete[] conv;
  }
};

struct Run {
  int kpl;        // planes per run
  size_t pbytes;  // device bytes of one plane's interior
};

Run plan_runs(const c3d_ctx* c) {
  Run r;
  r.pbytes = (size_t)c->g.nx * c->g.ny * c->esize;
  r.kpl = (int)std::max<size_t>(1, std::min<size_t>(kRunBytes / r.pbytes, (size_t)c->g.nz));
  return r;
}

int alloc_pinned(c3d_ctx* c, const Run& r, Pinned& pb) {
  for (int b = 0; b < 2; ++b) {
    C3D_HIP(hipHostMalloc(&pb.p[b], r.pbytes * r.kpl, hipHostMallocDefault));
    C3D_HIP(hipEventCreateWithFlags(&pb.ev[b], hipEventDisableTiming));
  }
  if (c->prec != C3D_FP64) pb.conv = new double[(size_t)c->g.nx * c->g.ny * r.kpl];
  return C3D_OK;
}

// device interior of planes [k0, k0+n) <-> host buffer (plane after plane)
int copy_run(c3d_ctx* c, const Run& r, void* host, void* dev, int k0, int n, bool to_host) {
  const Geom& g = c->g;
  char* d = (char*)dev + (size_t)fidx(g, 0, 1, k0) * c->esize;
  const size_t spitch = (size_t)g.plane * c->esize;
  if (to_host)
    C3D_HIP(hipMemcpy2DAsync(host, r.pbytes, d, spitch, r.pbytes, n, hipMemcpyDeviceToHost,
                             c->stream));
  else
    C3D_HIP(hipMemcpy2DAsync(d, spitch, host, r.pbytes, r.pbytes, n, hipMemcpyHostToDevice,
                             c->stream));
  return C3D_OK;
}

int io_fail(const std::string& what, const char* path) {
  set_error(what + ": " + path);
  return C3D_ERR_IO;
}

// one field: device -> file; shift is added to every value (uCRF)
int write_field(c3d_ctx* c, const Run& r, Pinned& pb, FILE* f, const void* dev, double shift,
                const char* path) {
  const int nz = c->g.nz;
  const size_t nplane = (size_t)c->g.nx * c->g.ny;
  int b = 0;
  int rc = copy_run(c, r, pb.p[0], (void*)dev, 0, std::min(r.kpl, nz), true);
  if (rc) return rc;
  C3D_HIP(hipEventRecord(pb.ev[0], c->stream));
  for (int k0 = 0; k0 < nz; k0 += r.kpl, b ^= 1) {
    const int n = std::min(r.kpl, nz - k0);
    const int k1 = k0 + r.kpl;
    if (k1 < nz) {
      rc = copy_run(c, r, pb.p[b ^ 1], (void*)dev, k1, std::min(r.kpl, nz - k1), true);
      if (rc) return rc;
      C3D_HIP(hipEventRecord(pb.ev[b ^ 1], c->stream));
    }
    C3D_HIP(hipEventSynchronize(pb.ev[b]));
    const size_t cnt = nplane * n;
    const double* out;
    if (c->prec == C3D_FP64) {
      double* h = (double*)pb.p[b];
      if (shift != 0.0)
        for (size_t q = 0; q < cnt; ++q) h[q] = h[q] + shift;
      out = h;
    } else {
      const float* h = (const float*)pb.p[b];
      for (size_t q = 0; q < cnt; ++q) pb.conv[q] = (double)h[q] + shift;
      out = pb.conv;
    }
    if (std::fwrite(out, sizeof(double), cnt, f) != cnt) {
      hipStreamSynchronize(c->stream);
      return io_fail("c3d_write_restart: write failed", path);
    }
  }
  return C3D_OK;
}

// one field: file -> device; shift is subtracted from every value (uCRF)
int read_field(c3d_ctx* c, const Run& r, Pinned& pb, FILE* f, void* dev, double shift,
               const char* path) {
  const int nz = c->g.nz;
  const size_t nplane = (size_t)c->g.nx * c->g.ny;
  int b = 0;
  for (int k0 = 0; k0 < nz; k0 += r.kpl, b ^= 1) {
    const int n = std::min(r.kpl, nz - k0);
    const size_t cnt = nplane * n;
    // the buffer's previous upload must have finished before it is refilled
    C3D_HIP(hipEventSynchronize(pb.ev[b]));
    if (c->prec == C3D_FP64) {
      double* h = (double*)pb.p[b];
      if (std::fread(h, sizeof(double), cnt, f) != cnt)
        return io_fail("c3d_read_restart: short read", path);
      if (shift != 0.0)
        for (size_t q = 0; q < cnt; ++q) h[q] = h[q] - shift;
    } else {
      if (std::fread(pb.conv, sizeof(double), cnt, f) != cnt)
        return io_fail("c3d_read_restart: short read", path);
      float* h = (float*)pb.p[b];
      for (size_t q = 0; q < cnt; ++q) h[q] = (float)(pb.conv[q] - shift);
    }
    int rc = copy_run(c, r, pb.p[b], dev, k0, n, false);
    if (rc) return rc;
    C3D_HIP(hipEventRecord(pb.ev[b], c->stream));
  }
  return C3D_OK;
}

int check_fields(const char* who, c3d_ctx* c, const char* path, const void* const* f, int nf) {
  if (!c || !path) {
    set_error(std::string(who) + ": bad arguments");
    return C3D_ERR_ARG;
  }
  for (int q = 0; q < 4; ++q)
    if (!f[q]) {
      set_error(std::string(who) + ": null velocity or pressure field");
      return C3D_ERR_ARG;
    }
  const int nh = (f[4] != nullptr) + (f[5] != nullptr) + (f[6] != nullptr);
  if (nh != 0 && nh != 3) {
    set_error(std::string(who) + ": give all three history fields or none");
    return C3D_ERR_ARG;
  }
  (void)nf;
  return C3D_OK;
}

}  // namespace

}  // namespace c3d

using namespace c3d;

extern "C" int c3d_restart_name(const char* restart_dir, const char* run_name, int ntime,
                                char* out, size_t out_len) {
  if (!restart_dir || !run_name || !out || ntime < 0 || ntime > 2147483647) {
    set_error("c3d_restart_name: bad arguments");
    return C3D_ERR_ARG;
  }
  const int n = std::snprintf(out, out_len, "%sRestartFor%s%010d", restart_dir, run_name, ntime);
  if (n < 0 || (size_t)n >= out_len) {
    set_error("c3d_restart_name: buffer too small");
    return C3D_ERR_ARG;
  }
  return C3D_OK;
}

extern "C" int c3d_write_restart(c3d_ctx* c, const char* path, const void* ux, const void* uy,
                                 const void* uz, const void* pr, const void* histx,
                                 const void* histy, const void* histz) {
  const void* f[7] = {ux, uy, uz, pr, histx, histy, histz};
  int rc = check_fields("c3d_write_restart", c, path, f, 7);
  if (rc) return rc;
  const int nf = histx ? 7 : 4;
  const Run r = plan_runs(c);
  Pinned pb(c->stream);
  if ((rc = alloc_pinned(c, r, pb))) return rc;
  C3D_HIP(hipStreamSynchronize(c->stream));
  FILE* fp = std::fopen(path, "wb");
  if (!fp) return io_fail("c3d_write_restart: cannot open file", path);
  const double ucrf = c->order == 4 ? c->p.u_crf : 0.0;
  for (int q = 0; q < nf && !rc; ++q)
    rc = write_field(c, r, pb, fp, f[q], q == 0 ? ucrf : 0.0, path);
  const bool closed = std::fclose(fp) == 0;
  if (rc) return rc;
  if (!closed) return io_fail("c3d_write_restart: write failed", path);
  return C3D_OK;
}

extern "C" int c3d_read_restart(c3d_ctx* c, const char* path, void* ux, void* uy, void* uz,
                                void* pr, void* histx, void* histy, void* histz) {
  void* f[7] = {ux, uy, uz, pr, histx, histy, histz};
  int rc = check_fields("c3d_read_restart", c, path, (const void* const*)f, 7);
  if (rc) return rc;
  const bool want_hist = histx != nullptr;
  struct stat st;
  if (::stat(path, &st) != 0) return io_fail("c3d_read_restart: cannot open file", path);
  const long long N = (long long)c->g.nx * c->g.ny * c->g.nz;
  const long long b7 = 8LL * N * 7, b4 = 8LL * N * 4, sz = (long long)st.st_size;
  // m_IOAndVisu.f90:455-467
  if (want_hist && sz != b7) return io_fail("c3d_read_restart: file byte wrong1", path);
  if (!want_hist && sz != b7 && sz != b4)
    return io_fail("c3d_read_restart: file byte wrong2", path);
  const Run r = plan_runs(c);
  Pinned pb(c->stream);
  if ((rc = alloc_pinned(c, r, pb))) return rc;
  FILE* fp = std::fopen(path, "rb");
  if (!fp) return io_fail("c3d_read_restart: cannot open file", path);
  C3D_HIP(hipStreamSynchronize(c->stream));
  // a new state velocity: the CFL of the next step must be computed from it
  for (int q = 0; q < 3; ++q)
    if (f[q] == c->state[C3D_UX] || f[q] == c->state[C3D_UY] || f[q] == c->state[C3D_UZ])
      c->cfl_valid = false;
  const double ucrf = c->order == 4 ? c->p.u_crf : 0.0;
  const int nf = want_hist ? 7 : 4;
  for (int q = 0; q < nf && !rc; ++q) rc = read_field(c, r, pb, fp, f[q], q == 0 ? ucrf : 0.0, path);
  std::fclose(fp);
  // the pinned buffers are freed on return: every upload must be complete
  hipError_t e = hipStreamSynchronize(c->stream);
  if (rc) return rc;
  if (e != hipSuccess) return hip_fail(e, "c3d_read_restart");
  return C3D_OK;
}
