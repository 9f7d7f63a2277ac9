// bmg_capi.hip -- the C ABI (include/bmagwa_hip.h): datasets, the drop-in
// Model, chain batches and their segment loop.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstddef>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "bmg_chain.h"
#include "bmg_device.h"
#include "bmg_internal.h"
#include "bmg_learn.h"
#include "bmg_rb.h"

using namespace bmg;

// kernels of bmg_chain.hip and bmg_data.hip
template <int KMAX> __global__ void k_chains(KernelArgs);
template <int KMAX> __global__ void k_model_op(ModelOpArgs);
__global__ void k_restride(const uint8_t*, uint64_t, uint32_t, uint32_t, uint8_t*);

static thread_local std::string g_err;

namespace bmg {
void set_error(const std::string& msg) { g_err = msg; }
int hip_fail(hipError_t e, const char* what) {
  g_err = std::string(what) + ": " + hipGetErrorString(e);
  return BMG_E_HIP;
}
}  // namespace bmg

extern "C" const char* bmg_last_error(void) { return g_err.c_str(); }

extern "C" int bmg_device_count(int* count) {
  BMG_HIP(hipGetDeviceCount(count));
  return BMG_OK;
}

// --------------------------------------------------------------- context --
static int init_types(bmg_ctx* c, const int32_t* types, int n_types) {
  if (n_types < 1 || n_types > 5) { set_error("n_types must be 1..5"); return BMG_E_ARG; }
  c->n_types = n_types;
  for (int i = 0; i < n_types; ++i) {
    int t = types[i];
    if (t < 0 || t > 4) { set_error("unknown effect type"); return BMG_E_ARG; }
    c->types[i] = t;
    c->allow_types[t] = 1;
    if (t == BMG_AH) { c->allow_terms[0] = 1; c->allow_terms[1] = 1; }
    else c->allow_terms[t] = 1;
  }
  // src/precomputed_snp_covariances.hpp:58-83
  int ah = c->allow_types[BMG_AH];
  int both = c->allow_types[BMG_A] && c->allow_types[BMG_H];
  int either = c->allow_types[BMG_A] || c->allow_types[BMG_H];
  c->offset = 2 * n_types + (ah ? 1 : 0) * (1 + (both ? -2 : 0) + (either ? 0 : 2));
  int k = 0;
  for (int j = 0; j < 4; ++j) c->offset_type[j] = c->allow_terms[j] ? k++ : -1;
  return BMG_OK;
}

static int ctx_alloc(int device, uint32_t n, uint64_t p, bmg_ctx** out, bmg_ctx** cp) {
  if (n < 2 || p < 1) { set_error("need n >= 2 and p >= 1"); return BMG_E_ARG; }
  BMG_HIP(hipSetDevice(device));
  bmg_ctx* c = new bmg_ctx();
  c->device = device;
  c->n = n; c->p = p;
  // rows of whole 128-byte lines: a scan step reads each SNP's line in one
  // piece instead of straddling two
  c->W32 = ((n + 15) / 16 + 31) / 32 * 32;
  c->n_pad = c->W32 * 16;
  hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e != hipSuccess) { delete c; return hip_fail(e, "hipStreamCreate"); }
  e = hipMalloc(&c->g, (size_t)p * c->W32 * 4);
  if (e != hipSuccess) { delete c; return hip_fail(e, "hipMalloc(genotypes)"); }
  e = hipMalloc(&c->y, sizeof(double) * c->n_pad);
  if (e != hipSuccess) { hipFree(c->g); delete c; return hip_fail(e, "hipMalloc(y)"); }
  hipMemset(c->y, 0, sizeof(double) * c->n_pad);
  *cp = c;
  *out = c;
  return BMG_OK;
}

extern "C" int bmg_ctx_set_y(bmg_ctx* c, const double* y) {
  if (!c || !y) { set_error("null argument"); return BMG_E_ARG; }
  c->y_host.assign(y, y + c->n);
  // src/data.hpp:67-69, in the reference's summation order
  double yy = 0.0, s = 0.0, sq = 0.0;
  for (uint32_t i = 0; i < c->n; ++i) { yy += y[i] * y[i]; }
  for (uint32_t i = 0; i < c->n; ++i) { sq += y[i] * y[i]; s += y[i]; }
  c->yy = yy;
  c->var_y = (sq - s * s / c->n) / (c->n - 1);
  BMG_HIP(hipMemcpy(c->y, y, sizeof(double) * c->n, hipMemcpyHostToDevice));
  return BMG_OK;
}

extern "C" int bmg_ctx_set_covariates(bmg_ctx* c, const double* e, uint32_t m_e) {
  if (!c || (m_e > 0 && !e)) { set_error("null argument"); return BMG_E_ARG; }
  if (m_e > 32) { set_error("at most 32 covariates on the device"); return BMG_E_UNSUPPORTED; }
  BMG_HIP(hipSetDevice(c->device));
  if (c->E) { BMG_HIP(hipFree(c->E)); c->E = nullptr; }
  c->m_e = m_e;
  c->e_host.assign(e, e + (size_t)c->n * m_e);
  if (m_e == 0) return BMG_OK;
  std::vector<double> padded((size_t)m_e * c->n_pad, 0.0);
  for (uint32_t k = 0; k < m_e; ++k)
    for (uint32_t i = 0; i < c->n; ++i) padded[(size_t)k * c->n_pad + i] = e[(size_t)k * c->n + i];
  BMG_HIP(hipMalloc(&c->E, sizeof(double) * padded.size()));
  BMG_HIP(hipMemcpy(c->E, padded.data(), sizeof(double) * padded.size(), hipMemcpyHostToDevice));
  return BMG_OK;
}

extern "C" int bmg_ctx_create(int device, const uint8_t* bed_body, uint32_t n, uint64_t p,
                              const double* y, const int32_t* types, int n_types, int recode,
                              bmg_ctx** out) {
  if (!bed_body || !y || !types || !out) { set_error("null argument"); return BMG_E_ARG; }
  bmg_ctx* c = nullptr;
  int rc = ctx_alloc(device, n, p, out, &c);
  if (rc) return rc;
  if ((rc = init_types(c, types, n_types))) { bmg_ctx_destroy(c); *out = nullptr; return rc; }
  // upload in the reference layout, then re-stride on the device
  uint32_t stride_raw = (n + 3) / 4;
  uint8_t* raw = nullptr;
  hipError_t e = hipMalloc(&raw, (size_t)stride_raw * p);
  if (e != hipSuccess) { bmg_ctx_destroy(c); *out = nullptr; return hip_fail(e, "hipMalloc(raw)"); }
  e = hipMemcpy(raw, bed_body, (size_t)stride_raw * p, hipMemcpyHostToDevice);
  if (e == hipSuccess) {
    unsigned gy = (unsigned)std::min<uint64_t>(p, 65535), gz = (unsigned)((p + gy - 1) / gy);
    hipLaunchKernelGGL(k_restride, dim3((c->W32 * 4 + 255) / 256, gy, gz), dim3(256), 0, c->stream, raw,
                       p, stride_raw, c->W32, (uint8_t*)c->g);
    e = hipStreamSynchronize(c->stream);
  }
  hipFree(raw);
  if (e != hipSuccess) { bmg_ctx_destroy(c); *out = nullptr; return hip_fail(e, "restride"); }
  if ((rc = ctx_prepare(c, recode))) { bmg_ctx_destroy(c); *out = nullptr; return rc; }
  if ((rc = bmg_ctx_set_y(c, y))) { bmg_ctx_destroy(c); *out = nullptr; return rc; }
  return BMG_OK;
}

extern "C" int bmg_ctx_create_synthetic(int device, uint32_t seed, uint32_t n, uint64_t p,
                                        const double* thresh, const double* rho,
                                        const int32_t* types, int n_types, bmg_ctx** out) {
  if (!thresh || !rho || !types || !out) { set_error("null argument"); return BMG_E_ARG; }
  bmg_ctx* c = nullptr;
  int rc = ctx_alloc(device, n, p, out, &c);
  if (rc) return rc;
  if ((rc = init_types(c, types, n_types))) { bmg_ctx_destroy(c); *out = nullptr; return rc; }
  std::vector<uint64_t> starts;
  for (uint64_t j = 0; j < p; ++j) if (j == 0 || rho[j] == 0.0) starts.push_back(j);
  double *dt = nullptr, *dr = nullptr; uint64_t* ds = nullptr;
  BMG_HIP(hipMalloc(&dt, sizeof(double) * p));
  BMG_HIP(hipMalloc(&dr, sizeof(double) * p));
  BMG_HIP(hipMalloc(&ds, sizeof(uint64_t) * starts.size()));
  BMG_HIP(hipMemcpy(dt, thresh, sizeof(double) * p, hipMemcpyHostToDevice));
  BMG_HIP(hipMemcpy(dr, rho, sizeof(double) * p, hipMemcpyHostToDevice));
  BMG_HIP(hipMemcpy(ds, starts.data(), sizeof(uint64_t) * starts.size(), hipMemcpyHostToDevice));
  rc = launch_synth(c, seed, dt, dr, ds, starts.size(), c->stream);
  hipError_t e = hipStreamSynchronize(c->stream);
  hipFree(dt); hipFree(dr); hipFree(ds);
  if (rc) { bmg_ctx_destroy(c); *out = nullptr; return rc; }
  if (e != hipSuccess) { bmg_ctx_destroy(c); *out = nullptr; return hip_fail(e, "synth"); }
  if ((rc = ctx_prepare(c, 0))) { bmg_ctx_destroy(c); *out = nullptr; return rc; }
  std::vector<double> zero(n, 0.0);
  c->y_host = zero;
  return BMG_OK;
}

extern "C" int bmg_ctx_create_subset(const bmg_ctx* pool, const uint32_t* individuals, uint32_t n_sub,
                                     const double* y, const int32_t* types, int n_types, bmg_ctx** out) {
  if (!pool || !individuals || !y || !types || !out) { set_error("null argument"); return BMG_E_ARG; }
  for (uint32_t i = 0; i < n_sub; ++i)
    if (individuals[i] >= pool->n) { set_error("individual index past the pool"); return BMG_E_ARG; }
  bmg_ctx* c = nullptr;
  int rc = ctx_alloc(pool->device, n_sub, pool->p, out, &c);
  if (rc) return rc;
  if ((rc = init_types(c, types, n_types))) { bmg_ctx_destroy(c); *out = nullptr; return rc; }
  uint32_t* d_sub = nullptr;
  hipError_t e = hipMalloc(&d_sub, sizeof(uint32_t) * std::max<uint32_t>(n_sub, 1));
  if (e != hipSuccess) { bmg_ctx_destroy(c); *out = nullptr; return hip_fail(e, "hipMalloc(subset)"); }
  e = hipMemcpy(d_sub, individuals, sizeof(uint32_t) * n_sub, hipMemcpyHostToDevice);
  // the pool's genotypes are final (its own stream has finished with them)
  if (e == hipSuccess) e = hipStreamSynchronize(pool->stream);
  if (e == hipSuccess) {
    rc = launch_subset(pool, d_sub, c, c->stream);
    e = hipStreamSynchronize(c->stream);
  }
  hipFree(d_sub);
  if (rc || e != hipSuccess) { bmg_ctx_destroy(c); *out = nullptr; return rc ? rc : hip_fail(e, "subset"); }
  if ((rc = ctx_prepare(c, 0))) { bmg_ctx_destroy(c); *out = nullptr; return rc; }
  if ((rc = bmg_ctx_set_y(c, y))) { bmg_ctx_destroy(c); *out = nullptr; return rc; }
  return BMG_OK;
}

extern "C" int bmg_ctx_destroy(bmg_ctx* c) {
  if (!c) return BMG_OK;
  hipSetDevice(c->device);
  if (c->stream) hipStreamSynchronize(c->stream);
  hipFree(c->g); hipFree(c->y); hipFree(c->pre); hipFree(c->E);
  hipFree(c->dig); hipFree(c->acc); hipFree(c->scale); hipFree(c->tmp_r); hipFree(c->tmp_out);
  if (c->stream) hipStreamDestroy(c->stream);
  delete c;
  return BMG_OK;
}

extern "C" int bmg_ctx_read_body(bmg_ctx* c, uint8_t* out) {
  if (!c || !out) { set_error("null argument"); return BMG_E_ARG; }
  uint32_t stride_raw = (c->n + 3) / 4;
  std::vector<uint8_t> buf((size_t)c->W32 * 4);
  for (uint64_t j = 0; j < c->p; ++j) {
    BMG_HIP(hipMemcpy(buf.data(), (uint8_t*)c->g + j * c->W32 * 4, c->W32 * 4, hipMemcpyDeviceToHost));
    std::memcpy(out + j * stride_raw, buf.data(), stride_raw);
  }
  return BMG_OK;
}

extern "C" int bmg_ctx_stats(bmg_ctx* c, double out[4]) {
  if (!c || !out) { set_error("null argument"); return BMG_E_ARG; }
  out[0] = c->yy; out[1] = c->var_y; out[2] = c->var_x; out[3] = c->mean_x;
  return BMG_OK;
}

extern "C" int bmg_ctx_missing_count(bmg_ctx* c, uint64_t* out) {
  if (!c || !out) { set_error("null argument"); return BMG_E_ARG; }
  *out = c->n_missing;
  return BMG_OK;
}

extern "C" int bmg_ctx_n_pad(bmg_ctx* c, uint32_t* n_pad) {
  if (!c || !n_pad) { set_error("null argument"); return BMG_E_ARG; }
  *n_pad = c->n_pad;
  return BMG_OK;
}

extern "C" int bmg_decode_many(bmg_ctx* c, const uint64_t* snps, uint64_t count, int coding,
                               double* out) {
  if (!c || !snps || !out || coding < 0 || coding > 3) { set_error("bad argument"); return BMG_E_ARG; }
  for (uint64_t i = 0; i < count; ++i)
    if (snps[i] >= c->p) { set_error("snp out of range"); return BMG_E_ARG; }
  BMG_HIP(hipSetDevice(c->device));
  uint64_t* ds = nullptr; double* dout = nullptr;
  BMG_HIP(hipMalloc(&ds, sizeof(uint64_t) * count));
  BMG_HIP(hipMalloc(&dout, sizeof(double) * count * c->n));
  BMG_HIP(hipMemcpy(ds, snps, sizeof(uint64_t) * count, hipMemcpyHostToDevice));
  int rc = launch_decode(c, ds, count, coding, dout, c->stream);
  if (!rc) {
    BMG_HIP(hipStreamSynchronize(c->stream));
    BMG_HIP(hipMemcpy(out, dout, sizeof(double) * count * c->n, hipMemcpyDeviceToHost));
  }
  hipFree(ds); hipFree(dout);
  return rc;
}

extern "C" int bmg_decode(bmg_ctx* c, uint64_t snp, int coding, double* out_n) {
  return bmg_decode_many(c, &snp, 1, coding, out_n);
}

extern "C" int bmg_self_moments_layout(bmg_ctx* c, int* offset, int offset_type[4]) {
  if (!c || !offset) { set_error("null argument"); return BMG_E_ARG; }
  *offset = c->offset;
  if (offset_type) for (int j = 0; j < 4; ++j) offset_type[j] = c->offset_type[j];
  return BMG_OK;
}

extern "C" int bmg_self_moments(bmg_ctx* c, double* out) {
  if (!c || !out) { set_error("null argument"); return BMG_E_ARG; }
  BMG_HIP(hipSetDevice(c->device));
  BMG_HIP(hipStreamSynchronize(c->stream));
  BMG_HIP(hipMemcpy(out, c->pre, sizeof(double) * c->offset * c->p, hipMemcpyDeviceToHost));
  return BMG_OK;
}

extern "C" int bmg_band_xtx_dev(bmg_ctx* c, uint32_t W, int ca, int cb, int32_t* out_dev, void* stream) {
  if (!c || !out_dev || ca < 0 || ca > 3 || cb < 0 || cb > 3) { set_error("bad argument"); return BMG_E_ARG; }
  hipStream_t s = stream ? (hipStream_t)stream : c->stream;
  return launch_band(c, W, ca, cb, out_dev, s);
}

extern "C" int bmg_band_xtx(bmg_ctx* c, uint32_t W, int ca, int cb, int32_t* out) {
  if (!c || !out) { set_error("null argument"); return BMG_E_ARG; }
  BMG_HIP(hipSetDevice(c->device));
  int32_t* d = nullptr;
  size_t bytes = sizeof(int32_t) * c->p * W;
  BMG_HIP(hipMalloc(&d, bytes));
  BMG_HIP(hipMemsetAsync(d, 0, bytes, c->stream));
  int rc = bmg_band_xtx_dev(c, W, ca, cb, d, nullptr);
  if (!rc) {
    BMG_HIP(hipStreamSynchronize(c->stream));
    BMG_HIP(hipMemcpy(out, d, bytes, hipMemcpyDeviceToHost));
  }
  hipFree(d);
  return rc;
}

extern "C" int bmg_scan_xtr_dev(bmg_ctx* c, const double* r_dev, int nvec, int coding, int digits,
                                double* out_dev, void* stream) {
  if (!c || !r_dev || !out_dev || coding < 0 || coding > 3) { set_error("bad argument"); return BMG_E_ARG; }
  hipStream_t s = stream ? (hipStream_t)stream : c->stream;
  return launch_scan(c, r_dev, nvec, coding, digits, out_dev, s);
}

extern "C" int bmg_scan_xtr_dev_noalloc(bmg_ctx* c, const double* r_dev, int nvec, int coding, int digits,
                                        double* out_dev, void* stream) {
  if (!c || !r_dev || !out_dev || coding < 0 || coding > 3) { set_error("bad argument"); return BMG_E_ARG; }
  hipStream_t s = stream ? (hipStream_t)stream : c->stream;
  return launch_scan_noalloc(c, r_dev, nvec, coding, digits, out_dev, s);
}

extern "C" int bmg_scan_xtr_timed(bmg_ctx* c, const double* r_dev, int nvec, int coding, int digits,
                                  double* out_dev, int reps, double* ms_per_launch, void* stream) {
  if (!c || !r_dev || !out_dev || !ms_per_launch || coding < 0 || coding > 3 || reps < 1) {
    set_error("bad argument"); return BMG_E_ARG;
  }
  hipStream_t s = stream ? (hipStream_t)stream : c->stream;
  return launch_scan_timed(c, r_dev, nvec, coding, digits, out_dev, reps, ms_per_launch, s);
}

extern "C" int bmg_scan_shape(bmg_ctx* c, int nvec, int digits, int32_t* out) {
  if (!c || !out || nvec < 1 || digits < 1 || nvec * digits > 16) { set_error("bad argument"); return BMG_E_ARG; }
  const int NC = nvec * digits;
  out[0] = c->scan_nt[NC];
  out[1] = (int32_t)c->scan_tw[NC];
  out[2] = c->scan_tw[NC] ? (int32_t)((c->W32 + c->scan_tw[NC] - 1) / c->scan_tw[NC]) : 0;
  out[3] = (int32_t)c->scan_gx[NC];
  out[4] = c->scan_wv[NC];
  return BMG_OK;
}

extern "C" int bmg_scan_set_shape(bmg_ctx* c, int nvec, int digits, int nt, int waves, uint32_t tile_words,
                                  uint32_t workgroups) {
  if (!c || nvec < 1 || digits < 1 || nvec * digits > 16 || (nt != 2 && nt != 4) ||
      (waves != 4 && waves != 8 && waves != 16) || tile_words < 32 || tile_words % 32 || workgroups < 1 ||
      ((size_t)std::min<uint32_t>(tile_words, (c->W32 + 31) / 32 * 32) * 16 + 16) * nvec * digits > 163840) {
    set_error("bad scan shape"); return BMG_E_ARG;
  }
  const int NC = nvec * digits;
  c->scan_nt[NC] = nt;
  c->scan_wv[NC] = waves;
  c->scan_tw[NC] = tile_words;
  c->scan_gx[NC] = workgroups;
  return BMG_OK;
}

extern "C" int bmg_band_xtx_timed(bmg_ctx* c, uint32_t W, int ca, int cb, int32_t* out_dev, int reps,
                                  double* ms_per_launch, void* stream) {
  if (!c || !out_dev || !ms_per_launch || ca < 0 || ca > 3 || cb < 0 || cb > 3 || reps < 1) {
    set_error("bad argument"); return BMG_E_ARG;
  }
  BMG_HIP(hipSetDevice(c->device));
  hipStream_t s = stream ? (hipStream_t)stream : c->stream;
  return launch_band_timed(c, W, ca, cb, out_dev, reps, ms_per_launch, s);
}

extern "C" int bmg_dev_alloc(int device, uint64_t bytes, void** ptr) {
  if (!ptr || bytes == 0) { set_error("bad argument"); return BMG_E_ARG; }
  *ptr = nullptr;
  BMG_HIP(hipSetDevice(device));
  BMG_HIP(hipMalloc(ptr, bytes));
  return BMG_OK;
}

extern "C" int bmg_dev_free(int device, void* ptr) {
  BMG_HIP(hipSetDevice(device));
  if (ptr) BMG_HIP(hipFree(ptr));
  return BMG_OK;
}

extern "C" int bmg_memcpy_h2d(int device, void* dst, const void* src, uint64_t bytes) {
  if (!dst || !src) { set_error("null argument"); return BMG_E_ARG; }
  BMG_HIP(hipSetDevice(device));
  BMG_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
  return BMG_OK;
}

extern "C" int bmg_memcpy_d2h(int device, void* dst, const void* src, uint64_t bytes) {
  if (!dst || !src) { set_error("null argument"); return BMG_E_ARG; }
  BMG_HIP(hipSetDevice(device));
  // the engine's streams do not synchronise with the null stream: wait for
  // every queued kernel before reading
  BMG_HIP(hipDeviceSynchronize());
  BMG_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
  return BMG_OK;
}

extern "C" int bmg_copy_timed(int device, const void* src_dev, void* dst_dev, uint64_t bytes, int reps,
                              double* ms_per_launch) {
  if (!src_dev || !dst_dev || !ms_per_launch || reps < 1) { set_error("bad argument"); return BMG_E_ARG; }
  BMG_HIP(hipSetDevice(device));
  hipStream_t s = nullptr;
  BMG_HIP(hipStreamCreate(&s));
  int rc = launch_copy_timed(src_dev, dst_dev, bytes, reps, ms_per_launch, s);
  hipStreamSynchronize(s);
  hipStreamDestroy(s);
  return rc;
}

extern "C" int bmg_read_timed(int device, const void* src_dev, uint64_t bytes, void* sink_dev, int reps,
                              double* ms_per_launch) {
  if (!src_dev || !sink_dev || !ms_per_launch || reps < 1) { set_error("bad argument"); return BMG_E_ARG; }
  BMG_HIP(hipSetDevice(device));
  hipStream_t s = nullptr;
  BMG_HIP(hipStreamCreate(&s));
  int rc = launch_read_timed(src_dev, bytes, sink_dev, reps, ms_per_launch, s);
  hipStreamSynchronize(s);
  hipStreamDestroy(s);
  return rc;
}

extern "C" int bmg_scan_xtr(bmg_ctx* c, const double* r, int nvec, int coding, int digits, double* out) {
  if (!c || !r || !out) { set_error("null argument"); return BMG_E_ARG; }
  BMG_HIP(hipSetDevice(c->device));
  if (ensure((void**)&c->tmp_r, &c->tmp_r_bytes, sizeof(double) * nvec * c->n_pad)) return BMG_E_HIP;
  if (ensure((void**)&c->tmp_out, &c->tmp_out_bytes, sizeof(double) * nvec * c->p)) return BMG_E_HIP;
  BMG_HIP(hipMemsetAsync(c->tmp_r, 0, sizeof(double) * nvec * c->n_pad, c->stream));
  for (int v = 0; v < nvec; ++v)
    BMG_HIP(hipMemcpyAsync(c->tmp_r + (size_t)v * c->n_pad, r + (size_t)v * c->n, sizeof(double) * c->n,
                           hipMemcpyHostToDevice, c->stream));
  int rc = launch_scan(c, c->tmp_r, nvec, coding, digits, c->tmp_out, c->stream);
  if (rc) return rc;
  BMG_HIP(hipStreamSynchronize(c->stream));
  BMG_HIP(hipMemcpy(out, c->tmp_out, sizeof(double) * nvec * c->p, hipMemcpyDeviceToHost));
  return BMG_OK;
}

// ----------------------------------------------------------------- batch --
struct bmg_batch {
  std::vector<bmg_ctx*> ctxs;
  int n_chains = 0;
  int kmax = 128;
  bmg_batch_cfg cfg{};
  std::vector<DevData> h_data;
  std::vector<ChainConst> h_consts;
  std::vector<ChainArrays> h_arrays;
  std::vector<int> chain_ds;
  DevData* d_data = nullptr;
  ChainConst* d_consts = nullptr;
  ChainScal* d_scal = nullptr;
  ChainArrays* d_arrays = nullptr;
  uint8_t* pool = nullptr;
  size_t per_chain = 0;
  uint32_t n_rao = 0, thin = 0, n_st = 0;
  uint64_t iter = 0;
  int setup_done = 0;
  int device = 0;
  hipStream_t stream = nullptr;
  // RB scratch
  double* r = nullptr; size_t r_bytes = 0;
  double* xa = nullptr; size_t xa_bytes = 0;
  double* xh = nullptr; size_t xh_bytes = 0;
  double* rsum = nullptr;
  int* d_list = nullptr;
  // timing of the last run
  double t_total = 0, t_chain = 0, t_scan = 0;
  int n_scans = 0;
  // graph of one steady-state segment (n_rao iterations + its RB block)
  hipGraphExec_t graph_exec = nullptr;
  hipGraph_t graph = nullptr;
  int* d_glist = nullptr;                 // every dataset's chains, in dataset order
  std::vector<const void*> graph_sig;     // buffers the graph's nodes point at
  int graph_scans = 0;                    // scan launches per segment
  double n_replays = 0, t_graph = 0, n_builds = 0;
  // learned-swap weights per dataset (p x 2H floats)
  std::vector<float*> lw_dev;
  // online learned proposal per dataset (proposal = 1)
  std::vector<LearnDS*> learn;
  std::vector<int> learn_on;
  // x_j . y tables per dataset and coding (set-up)
  std::vector<double*> xy_dev;
};

struct bmg_model {
  bmg_batch* b = nullptr;
};

static void tree_layout(DevData& D) {
  uint64_t len = (D.p + 63) / 64 * 64;
  int h = 0;
  uint64_t off = 0;
  for (;;) {
    D.lev_len[h] = len;
    D.lev_off[h] = off;
    off += len;
    if (len == 64) break;
    len = ((len / 64) + 63) / 64 * 64;
    ++h;
  }
  D.top = h;
  D.tree_size = off;
  D.lds_lev_size = D.top >= 2 ? off - D.lev_off[2] : 0;
}

static DevData make_devdata(bmg_ctx* c) {
  DevData D{};
  D.g = c->g; D.y = c->y; D.pre = c->pre;
  D.p = c->p; D.n = c->n; D.W32 = c->W32; D.n_pad = c->n_pad;
  D.offset = c->offset;
  for (int j = 0; j < 4; ++j) D.ot[j] = c->offset_type[j];
  D.type = c->types[0];
  D.n_types = c->n_types;
  for (int i = 0; i < 5; ++i) D.types[i] = c->types[i];
  D.me = 1 + (int)c->m_e;
  D.E = c->E;
  tree_layout(D);
  return D;
}

// src/prior.hpp:273-290
static int solve_beta(double m_g, double e_q, double var_q, double* ga, double* gb) {
  if (m_g > 1) {
    double z = (var_q - e_q * (1 - e_q)) / ((m_g - 1) * e_q);
    *ga = (z - 1) / (1 - m_g * z / e_q);
    *gb = (m_g / e_q - 1) * *ga;
  } else {
    double z = e_q / (1 - e_q);
    *gb = z / var_q / std::pow(1 + z, 3) - 1 / (1 + z);
    *ga = z * *gb;
  }
  if (!(*ga > 0) || !(*gb > 0)) return -1;
  return 0;
}

static void geometric_cdf_h(int maxsize, double p, double* values) {
  double q = 1 - p, qp = q;
  for (int i = 0; i < maxsize; ++i) { values[i] = 1 - q; q *= qp; }
}

extern "C" __global__ void k_init_chain_arrays(const DevData* data, const ChainConst* consts,
                                              const ChainArrays* arrays, const double* p0) {
  int c = blockIdx.y;
  const ChainConst& K = consts[c];
  const DevData& D = data[K.ds];
  const ChainArrays& A = arrays[c];
  for (uint64_t j = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; j < D.lev_len[0];
       j += (uint64_t)gridDim.x * blockDim.x) {
    if (j < D.p) {
      double pp = p0[c];
      double qa = fmax(pp, K.q_add_min), qr = fmax(1 - pp, K.q_rem_min);
      A.model_inds[j] = -1;
      A.p_proposal[j] = pp; A.p_rao[j] = 0.0; A.p_r[j] = 0.0;
      A.q_add[j] = qa; A.q_rem[j] = qr;
      A.tree_add[j] = qa;     // model empty: nothing zeroed for additions
      A.tree_rem[j] = 0.0;    // ... and everything zeroed for removals
    } else {
      A.tree_add[j] = 0.0; A.tree_rem[j] = 0.0;
    }
  }
}

// Below 256 columns an iterating segment is two launches on the stream: the
// batch capacity's kernel, which hands a chain whose model comes within one
// iteration's growth of its capacity to the 256-column kernel at that
// iteration, then the 256-column kernel, which runs those chains to the end
// of the segment (its other workgroups return at once).
static int launch_chain_kernel(bmg_batch* b, uint64_t n_iter, int mode) {
  const int two = mode == 0 && b->kmax < BMG_KMAXG;
  KernelArgs a{b->d_data, b->d_consts, b->d_scal, b->d_arrays, b->n_chains, n_iter, mode, two ? 1 : 0, 0};
  switch (b->kmax) {
    case 64: hipLaunchKernelGGL(k_chains<64>, dim3(b->n_chains), dim3(256), 0, b->stream, a); break;
    case 128: hipLaunchKernelGGL(k_chains<128>, dim3(b->n_chains), dim3(256), 0, b->stream, a); break;
    case 160: hipLaunchKernelGGL(k_chains<160>, dim3(b->n_chains), dim3(256), 0, b->stream, a); break;
    default: hipLaunchKernelGGL(k_chains<256>, dim3(b->n_chains), dim3(256), 0, b->stream, a); break;
  }
  BMG_HIP(hipGetLastError());
  if (two) {
    KernelArgs a2 = a;
    a2.role = 2;
    a2.small_kmax = b->kmax;
    hipLaunchKernelGGL(k_chains<BMG_KMAXG>, dim3(b->n_chains), dim3(256), 0, b->stream, a2);
    BMG_HIP(hipGetLastError());
  }
  return BMG_OK;
}

static int ensure_rb_scratch(bmg_batch* b, int nv, uint32_t n_pad, uint64_t p) {
  if (ensure((void**)&b->r, &b->r_bytes, sizeof(double) * nv * n_pad)) return BMG_E_HIP;
  if (ensure((void**)&b->xa, &b->xa_bytes, sizeof(double) * nv * p)) return BMG_E_HIP;
  if (ensure((void**)&b->xh, &b->xh_bytes, sizeof(double) * nv * p)) return BMG_E_HIP;
  return BMG_OK;
}

// RB block for the chains in `list` (all on dataset ds)
static int rb_block_ds(bmg_batch* b, int ds, const std::vector<int>& chains, int learn_now) {
  bmg_ctx* c = b->ctxs[ds];
  const DevData& D = b->h_data[ds];
  int digits = b->cfg.scan_digits > 0 ? b->cfg.scan_digits : 5;
  int per = 16 / digits;
  for (size_t s0 = 0; s0 < chains.size(); s0 += per) {
    int nv = (int)std::min<size_t>(per, chains.size() - s0);
    if (ensure_rb_scratch(b, nv, D.n_pad, D.p)) return BMG_E_HIP;
    BMG_HIP(hipMemcpyAsync(b->d_list, chains.data() + s0, sizeof(int) * nv, hipMemcpyHostToDevice,
                           b->stream));
    if (launch_residuals(c->y, D.n, D.n_pad, b->d_arrays, b->d_list, nv, b->r, b->rsum, b->stream)) {
      set_error("residual kernel launch failed"); return BMG_E_HIP;
    }
    int t = D.type;
    int rc = launch_scan(c, b->r, nv, t == BMG_AH ? 0 : t, digits, b->xa, b->stream);
    if (rc) return rc;
    if (t == BMG_AH) {
      rc = launch_scan(c, b->r, nv, 1, digits, b->xh, b->stream);
      if (rc) return rc;
    }
    b->n_scans += (t == BMG_AH) ? 2 : 1;
    if (launch_rb_finalize(b->d_data, b->d_consts, b->d_scal, b->d_arrays, b->d_list, nv, BMG_KMAXG, D.p,
                           b->xa, b->xh, b->rsum, b->stream)) {
      set_error("rb finalize launch failed"); return BMG_E_HIP;
    }
    if (learn_now) {
      rc = learn_labels(b->learn[ds], b->d_data, b->d_consts, b->d_scal, b->d_arrays, b->d_list, nv, BMG_KMAXG,
                        b->xa, b->rsum, b->stream);
      if (!rc) rc = learn_snp_labels(b->learn[ds], b->d_scal, b->d_arrays, b->d_list, nv, BMG_KMAXG, b->stream);
      if (rc) return rc;
    }
  }
  return BMG_OK;
}

// the full RB block after an RB iteration: scans, element-wise updates,
// tree rebuilds, clearing the pending flags
static int set_learn_on(bmg_batch* b, int ds, int on, double p_learned);

// learned proposal after an RB block of dataset ds: a training round every
// learn_every iterations of burn-in and at its last block, which freezes the
// table (no sample after burn-in sees it change)
static int learn_after_block(bmg_batch* b, int ds, const std::vector<int>& all,
                             const std::vector<ChainScal>& cs) {
  LearnDS* L = b->learn[ds];
  int burn = 0, last = 1;
  for (int c : all) {
    if (!cs[c].rb_accum_rao) burn = 1;        // this block was a burn-in block
    if (cs[c].n_rao_burnin > 0) last = 0;     // more burn-in blocks follow
  }
  if (!burn) {
    if (!learn_frozen(L)) learn_freeze(L);    // burn-in already over: never train
    return BMG_OK;
  }
  const uint64_t every = (uint64_t)std::max(b->cfg.learn_every, 1);
  int freeze = last || (b->cfg.learn_freeze_iter > 0 && b->iter >= (uint64_t)b->cfg.learn_freeze_iter);
  if (!freeze && b->iter % every != 0 && b->iter / every == (b->iter - b->n_rao) / every) return BMG_OK;
  std::vector<int> chains;
  for (int c = 0; c < b->n_chains; ++c) if (b->chain_ds[c] == ds) chains.push_back(c);
  int rc = learn_round(L, b->d_arrays, chains.data(), (int)chains.size(), b->stream);
  if (rc) return rc;
  if ((b->cfg.proposal & 1) && learn_rounds(L) > 0 && !b->learn_on[ds]) {
    rc = set_learn_on(b, ds, 1, b->cfg.p_learned);
    if (rc) return rc;
    b->learn_on[ds] = 1;
  }
  if (freeze && (b->cfg.proposal & 2)) {
    // the per-SNP network's tables replace the chains' proposals, after
    // this block's adaptation step (the last of burn-in at the default
    // freeze point)
    const DevData& D = b->h_data[ds];
    rc = learn_snp_apply(L, b->d_data, b->d_consts, b->d_scal, b->d_arrays, chains.data(), (int)chains.size(),
                         D.lev_len, D.top, b->stream);
    if (rc) return rc;
  }
  if (freeze) learn_freeze(L);
  BMG_HIP(hipStreamSynchronize(b->stream));
  return BMG_OK;
}

// a chain whose current model became non-finite has stopped iterating: the
// run fails with the chain's index (the oracle stops at the same point)
static int numeric_error(const std::vector<ChainScal>& cs) {
  for (size_t c = 0; c < cs.size(); ++c)
    if (cs[c].err & ERR_TRACE) {
      set_error("chain " + std::to_string(c) + ": trace buffer full (raise trace_capacity or do_n_iter)");
      return BMG_E_CAPACITY;
    }
  for (size_t c = 0; c < cs.size(); ++c)
    if (cs[c].err & ERR_CAPACITY) {
      // the reference stops the same way at its compiled limit
      // (src/model.hpp:247-262, MAXMODELSIZE)
      set_error("chain " + std::to_string(c) + " stopped at iteration " + std::to_string(cs[c].n_iter) +
                ": a proposal needed more than " + std::to_string(BMG_KMAXG) +
                " model columns, the device limit (set max_model_size)");
      return BMG_E_CAPACITY;
    }
  for (size_t c = 0; c < cs.size(); ++c)
    if (cs[c].err & ERR_STOP) {
      const ChainScal& x = cs[c];
      char buf[384];
      snprintf(buf, sizeof(buf),
               "chain %zu stopped at iteration %llu (flags %d:%s%s%s%s): loglik %.6g syx %.6g sigma2 %.6g "
               "alpha %.6g columns %d loci %d",
               c, (unsigned long long)x.n_iter, x.err, (x.err & ERR_NONFINITE) ? " refresh" : "",
               (x.err & ERR_SIGMA2) ? " sigma2" : "", (x.err & ERR_ALPHA) ? " alpha" : "",
               (x.err & ERR_WALK) ? " walk" : "", x.ll, x.syx, x.sigma2, x.alpha, x.cols, x.nloci);
      set_error(buf);
      return BMG_E_NUMERIC;
    }
  return BMG_OK;
}

static int rb_block(bmg_batch* b) {
  std::vector<ChainScal> cs(b->n_chains);
  BMG_HIP(hipMemcpyAsync(cs.data(), b->d_scal, sizeof(ChainScal) * b->n_chains, hipMemcpyDeviceToHost,
                         b->stream));
  BMG_HIP(hipStreamSynchronize(b->stream));
  if (int rc = numeric_error(cs)) return rc;
  for (size_t ds = 0; ds < b->ctxs.size(); ++ds) {
    std::vector<int> scan_list, all;
    for (int c = 0; c < b->n_chains; ++c) {
      if (b->chain_ds[c] != (int)ds || !cs[c].rb_pending) continue;
      all.push_back(c);
      if (cs[c].rb_do_scan) scan_list.push_back(c);
    }
    int learn_now = 0;
    if (ds < b->learn.size() && b->learn[ds] && !learn_frozen(b->learn[ds]))
      for (int c : scan_list) if (!cs[c].rb_accum_rao) learn_now = 1;
    if (!scan_list.empty()) {
      int rc = rb_block_ds(b, (int)ds, scan_list, learn_now);
      if (rc) return rc;
    }
    if (!all.empty()) {
      const DevData& D = b->h_data[ds];
      for (size_t s0 = 0; s0 < all.size(); s0 += 1024) {
        int nv = (int)std::min<size_t>(1024, all.size() - s0);
        BMG_HIP(hipMemcpyAsync(b->d_list, all.data() + s0, sizeof(int) * nv, hipMemcpyHostToDevice,
                               b->stream));
        if (launch_rb_update(b->d_data, b->d_consts, b->d_scal, b->d_arrays, b->d_list, nv, D.p,
                             b->stream) ||
            launch_tree_build(b->d_data, b->d_consts, b->d_scal, b->d_arrays, b->d_list, nv, D.lev_len,
                              D.top, 1, 1, b->stream)) {
          set_error("rb update launch failed"); return BMG_E_HIP;
        }
        BMG_HIP(hipStreamSynchronize(b->stream));
      }
      if (ds < b->learn.size() && b->learn[ds] && !learn_frozen(b->learn[ds])) {
        int rc = learn_after_block(b, (int)ds, all, cs);
        if (rc) return rc;
      }
    }
  }
  return BMG_OK;
}

extern "C" int bmg_batch_destroy(bmg_batch* b) {
  if (!b) return BMG_OK;
  hipSetDevice(b->device);
  if (b->stream) hipStreamSynchronize(b->stream);
  hipFree(b->pool); hipFree(b->d_data); hipFree(b->d_consts); hipFree(b->d_scal); hipFree(b->d_arrays);
  for (float* w : b->lw_dev) hipFree(w);
  for (LearnDS* L : b->learn) learn_destroy(L);
  hipFree(b->r); hipFree(b->xa); hipFree(b->xh); hipFree(b->rsum); hipFree(b->d_list);
  if (b->graph_exec) hipGraphExecDestroy(b->graph_exec);
  if (b->graph) hipGraphDestroy(b->graph);
  hipFree(b->d_glist);
  for (double* x : b->xy_dev) hipFree(x);
  if (b->stream) hipStreamDestroy(b->stream);
  delete b;
  return BMG_OK;
}

static size_t align256(size_t x) { return (x + 255) / 256 * 256; }

static int batch_create(bmg_ctx** ctxs, int n_ctx, int n_chains, const int32_t* chain_ctx,
                        const bmg_opts* opts, const bmg_batch_cfg* cfg, int model_api, bmg_batch** out) {
  if (!ctxs || n_ctx < 1 || n_chains < 1 || !chain_ctx || !opts || !out) {
    set_error("null or empty argument"); return BMG_E_ARG;
  }
  *out = nullptr;
  bmg_batch* b = new bmg_batch();
  b->ctxs.assign(ctxs, ctxs + n_ctx);
  b->n_chains = n_chains;
  if (cfg) b->cfg = *cfg;
  int kmax = 128;
  b->device = ctxs[0]->device;
  hipSetDevice(b->device);
  // capacity of the launch: the smallest compiled one that holds a typical
  // model (the prior mean plus one move, AH: two columns a locus); a chain
  // whose model comes within one move of it continues in the 256-column
  // kernel from that iteration, so the capacity changes speed, not results
  double typical = 0;
  if (b->cfg.max_model_size < 0) { delete b; set_error("max_model_size must be >= 0"); return BMG_E_ARG; }
  for (int c = 0; c < n_chains; ++c) {
    double typ = opts[c].e_qg + (double)opts[c].max_move_size + 8.0;
    if (b->cfg.max_model_size > 0)
      typ = std::min(typ, (double)b->cfg.max_model_size + (double)opts[c].max_move_size + 1.0);
    for (int t = 0; t < opts[c].n_types; ++t)
      if (opts[c].types[t] == BMG_AH) typ *= 2;
    typical = std::max(typical, typ + 1);
  }
  kmax = typical <= 64 ? 64 : (typical <= 128 ? 128 : (typical <= 160 ? 160 : 256));
  if (model_api) kmax = 160;
  if (b->cfg.capacity) {
    if (b->cfg.capacity != 64 && b->cfg.capacity != 128 && b->cfg.capacity != 160 && b->cfg.capacity != 256) {
      delete b; set_error("capacity must be 0, 64, 128, 160 or 256");
      return BMG_E_ARG;
    }
    if (!model_api) kmax = b->cfg.capacity;
  }
  b->kmax = kmax;
  uint64_t dr_max = 0;
  for (int c = 0; c < n_chains; ++c) {
    const bmg_opts& o = opts[c];
    if (chain_ctx[c] < 0 || chain_ctx[c] >= n_ctx) { delete b; set_error("bad chain_ctx"); return BMG_E_ARG; }
    bmg_ctx* x = ctxs[chain_ctx[c]];
    if (o.sampler_type != 0) { delete b; set_error("only the PMV sampler runs on the device"); return BMG_E_UNSUPPORTED; }
    if (!model_api && (o.n_types != 1 || x->n_types != 1 || o.types[0] != x->types[0])) {
      delete b; set_error("the device sampler needs exactly one effect type, equal to the dataset's");
      return BMG_E_UNSUPPORTED;
    }
    if (o.m_e != x->m_e) {
      delete b; set_error("m_e of the options differs from the dataset's covariates (bmg_ctx_set_covariates)");
      return BMG_E_ARG;
    }
    if (!model_api && x->n_missing != 0) { delete b; set_error("missing genotypes are not supported on the device"); return BMG_E_UNSUPPORTED; }
    if (o.thin == 0 || o.n_rao == 0 || o.n_sample_tau2_and_missing == 0 || o.n_rao % o.thin != 0) {
      delete b; set_error("need thin, n_rao, n_sample_tau2_and_missing > 0 and n_rao % thin == 0");
      return BMG_E_ARG;
    }
    if (c == 0) { b->n_rao = o.n_rao; b->thin = o.thin; b->n_st = o.n_sample_tau2_and_missing; }
    else if (o.n_rao != b->n_rao || o.thin != b->thin || o.n_sample_tau2_and_missing != b->n_st) {
      delete b; set_error("all chains of a batch share n_rao, thin and n_sample_tau2_and_missing");
      return BMG_E_ARG;
    }
    uint64_t mm = std::min<uint64_t>(std::min<uint64_t>(x->p, 255), o.max_move_size);
    if (mm > BMG_MAXMS || mm < 1) { delete b; set_error("max_move_size must be in 1..64 on the device"); return BMG_E_UNSUPPORTED; }
    // the second stage scores 2^dr sub-models (src/sampler.hpp:163,272-281)
    const uint64_t dr = std::min<uint64_t>(mm, o.delay_rejection);
    if (dr > BMG_DR_MAX) {
      delete b; set_error("min(delay_rejection, max_move_size) must be at most 20 on the device");
      return BMG_E_UNSUPPORTED;
    }
    dr_max = std::max(dr_max, dr);
  }
  for (int i = 0; i < n_ctx; ++i) {
    b->h_data.push_back(make_devdata(ctxs[i]));
    b->h_data.back().mixed = model_api ? 1 : 0;
    if (b->h_data.back().lds_lev_size > 512) { delete b; set_error("p too large for the sum-tree cache"); return BMG_E_UNSUPPORTED; }
  }
  b->chain_ds.assign(chain_ctx, chain_ctx + n_chains);
  BMG_HIP(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
  // per-chain layout
  // trace buffers sized for the longest chain; thinned loci for models of
  // up to BMG_KMAXG loci on average unless the caller sizes them
  uint64_t n_iter_max = 1;
  for (int c = 0; c < n_chains; ++c) n_iter_max = std::max<uint64_t>(n_iter_max, opts[c].do_n_iter);
  uint64_t tr_cap = b->cfg.trace_every_iter ? n_iter_max + 8 : 1;
  uint64_t th_cap = n_iter_max / opts[0].thin + 8;
  uint64_t th_loci_cap = b->cfg.trace_capacity ? b->cfg.trace_capacity : th_cap * BMG_KMAXG;
  uint64_t pmax = 0, npmax = 0, tmax = 0, wmax = 0;
  for (auto& D : b->h_data) {
    pmax = std::max(pmax, D.p); npmax = std::max<uint64_t>(npmax, D.n_pad); tmax = std::max(tmax, D.tree_size);
    wmax = std::max<uint64_t>(wmax, D.W32);
  }
  size_t sz = 0;
  auto add = [&](size_t bytes) { size_t o = sz; sz += align256(bytes); return o; };
  size_t o_mi = add(4 * pmax), o_qa = add(8 * pmax), o_qr = add(8 * pmax);
  size_t o_ta = add(8 * tmax), o_tr = add(8 * tmax);
  size_t o_pp = add(8 * pmax), o_pr = add(8 * pmax), o_prr = add(8 * pmax);
  // per-column arrays for BMG_KMAXG columns whatever the launch capacity K:
  // a chain that leaves K continues in the 256-column kernel on the same state
  const size_t G = BMG_KMAXG;
  size_t o_L = add(8 * G * (G + 1) / 2), o_xx = add(8 * G * G);
  size_t o_Lw = add(8 * G * (G + 1) / 2);
  size_t o_xc = add(4 * G * wmax);
  size_t o_yh = add(8 * npmax), o_vec = add(8 * 5 * G), o_idx = add(4 * 8 * G), o_ad = add(8 * 6 * 256);
  size_t o_trm = add(tr_cap), o_trj = add(tr_cap), o_trs = add(tr_cap), o_trz = add(4 * tr_cap),
         o_trl = add(8 * tr_cap);
  size_t o_ths = add(4 * th_cap), o_thl = add(4 * th_loci_cap), o_thll = add(8 * th_cap),
         o_ths2 = add(8 * th_cap), o_tha = add(8 * th_cap), o_thp = add(8 * 3 * th_cap);
  size_t o_dr = add(dr_max ? (size_t)8 << dr_max : 0);
  // save_beta: every thinned sample's beta (the fixed columns plus up to two
  // per locus)
  int any_beta = 0;
  for (int c = 0; c < n_chains; ++c) any_beta |= opts[c].save_beta != 0;
  const uint64_t th_beta_cap = any_beta ? 2 * th_loci_cap + th_cap * (uint64_t)(1 + 32) : 0;
  size_t o_thb = add(8 * th_beta_cap);
  b->per_chain = sz;
  hipError_t e = hipMalloc(&b->pool, sz * n_chains);
  if (e != hipSuccess) { bmg_batch_destroy(b); return hip_fail(e, "hipMalloc(chain pool)"); }
  BMG_HIP(hipMemsetAsync(b->pool, 0, sz * n_chains, b->stream));
  std::vector<ChainScal> scal(n_chains);
  std::vector<double> p0(n_chains);
  std::vector<double> vec(5 * G), adapt(6 * 256);
  std::vector<int32_t> idx(8 * G);
  for (int c = 0; c < n_chains; ++c) {
    const bmg_opts& o = opts[c];
    bmg_ctx* x = ctxs[chain_ctx[c]];
    uint8_t* base = b->pool + sz * c;
    ChainArrays A{};
    A.model_inds = (int32_t*)(base + o_mi); A.q_add = (double*)(base + o_qa); A.q_rem = (double*)(base + o_qr);
    A.tree_add = (double*)(base + o_ta); A.tree_rem = (double*)(base + o_tr);
    A.p_proposal = (double*)(base + o_pp); A.p_rao = (double*)(base + o_pr); A.p_r = (double*)(base + o_prr);
    A.Lcur = (double*)(base + o_L); A.xxs = (double*)(base + o_xx); A.y_hat = (double*)(base + o_yh);
    A.Lwork = (double*)(base + o_Lw);
    A.xcol = (uint32_t*)(base + o_xc);
    A.vec = (double*)(base + o_vec); A.idx = (int32_t*)(base + o_idx); A.adapt = (double*)(base + o_ad);
    A.tr_move = base + o_trm; A.tr_jd = base + o_trj; A.tr_ms = base + o_trs;
    A.tr_size = (uint32_t*)(base + o_trz); A.tr_ll = (double*)(base + o_trl);
    A.th_size = (uint32_t*)(base + o_ths); A.th_loci = (uint32_t*)(base + o_thl);
    A.th_ll = (double*)(base + o_thll); A.th_s2 = (double*)(base + o_ths2); A.th_alpha = (double*)(base + o_tha);
    A.th_pve = (double*)(base + o_thp);
    A.dr_lp = dr_max ? (double*)(base + o_dr) : nullptr;
    A.th_beta = o.save_beta ? (double*)(base + o_thb) : nullptr;
    b->h_arrays.push_back(A);
    // constants (src/sampler.hpp:107-203, src/prior.hpp:45-122)
    ChainConst Kc{};
    Kc.seed = o.seed; Kc.chain_id = o.chain; Kc.ds = (uint32_t)chain_ctx[c];
    Kc.thin = o.thin; Kc.n_rao = o.n_rao; Kc.n_st = o.n_sample_tau2_and_missing;
    Kc.max_ms = (uint32_t)std::min<uint64_t>(std::min<uint64_t>(x->p, 255), o.max_move_size);
    Kc.max_nbh = o.max_SNP_neighborhood_size;
    Kc.dr = (int32_t)std::min<uint64_t>(Kc.max_ms, o.delay_rejection);
    Kc.max_size = b->cfg.max_model_size > 0 ? b->cfg.max_model_size : INT32_MAX;
    // a move adds at most max_move_size loci (two columns each for AH);
    // one column of slack
    Kc.col_reserve = (int32_t)Kc.max_ms * (x->types[0] == BMG_AH ? 2 : 1) + 1;
    Kc.adapt_p_ms = o.adapt_p_move_size && Kc.max_ms > 1;
    Kc.adaptation = o.adaptation; Kc.flat = o.flat_proposal_dist;
    Kc.trace_iter = b->cfg.trace_every_iter;
    Kc.acpt_goal = o.p_move_size_acpt_goal;
    double m_g = (double)x->p;
    Kc.q_add_min = 1.0 / (m_g - o.e_qg);
    Kc.q_rem_min = 1.0 / o.e_qg;
    Kc.k_move_size = 1000.0 / (double)o.n_rao_burnin;
    double s2_sigma2 = o.s2_sigma2 > 0 ? o.s2_sigma2
        : x->var_y * (1 - o.R2mode_sigma2) * (o.nu_sigma2 + 2) / o.nu_sigma2;
    Kc.n_plus_nu = x->n + o.nu_sigma2;
    Kc.minus_n_plus_nu_2 = -0.5 * Kc.n_plus_nu;
    Kc.nus2_plus_yy = o.nu_sigma2 * s2_sigma2 + x->yy;
    Kc.nus2 = o.nu_sigma2 * s2_sigma2;
    Kc.mu_alpha = o.mu_alpha;
    double ga, gb;
    if (solve_beta(m_g, o.e_qg, o.var_qg, &ga, &gb)) { bmg_batch_destroy(b); set_error("invalid a or b of the model-size prior"); return BMG_E_ARG; }
    Kc.g_a = ga; Kc.g_b = gb; Kc.m_g = m_g;
    int type = x->types[0];
    Kc.tp_type = o.types_prior[type]; Kc.tp_sum = o.types_prior[type];
    double alpha = std::max(0.5, o.mu_alpha), alpha2 = alpha * alpha;
    ChainScal& S = scal[c];
    for (int t = 0; t < 4; ++t) {
      Kc.allow_terms[t] = x->allow_terms[t];
      if (!x->allow_terms[t]) { Kc.nu_tau2[t] = Kc.s2_tau2[t] = Kc.nus2_tau2[t] = NAN; S.inv_tau2_alpha2[t] = NAN; continue; }
      double s2t = o.s2_tau2[t] > 0 ? o.s2_tau2[t]
          : o.eh_tau2[t] * (o.nu_tau2[t] - 2) / (o.nu_tau2[t] * (x->var_x + x->mean_x * x->mean_x) * (1 - o.R2mode_sigma2));
      if (!(s2t > 0) || !(o.nu_tau2[t] > 0)) { bmg_batch_destroy(b); set_error("invalid tau2 prior"); return BMG_E_ARG; }
      Kc.nu_tau2[t] = o.nu_tau2[t]; Kc.s2_tau2[t] = s2t; Kc.nus2_tau2[t] = o.nu_tau2[t] * s2t;
      double tau2_init = Kc.nus2_tau2[t] / std::max(0.1, o.nu_tau2[t] - 2);
      S.inv_tau2_alpha2[t] = 1.0 / (alpha2 * tau2_init);
    }
    Kc.use_ind = o.use_individual_tau2;
    Kc.inv_tau2_e0 = o.inv_tau2_e_const_val;
    Kc.tr_cap = tr_cap; Kc.th_cap = th_cap; Kc.th_loci_cap = th_loci_cap;
    Kc.th_beta_cap = o.save_beta ? th_beta_cap : 0;
    b->h_consts.push_back(Kc);
    // initial scalars: empty model (constant column only)
    S.alpha = alpha; S.alpha2 = alpha2;
    S.n_rao_burnin = o.n_rao_burnin; S.p_rao_n = 0; S.p_proposal_n = 1;
    S.p_ms = o.p_move_size;
    for (int q = 0; q < 3; ++q) { S.n_acpt[q] = 1; S.n_moves[q] = 2; }
    const int me = 1 + (int)x->m_e;   // the constant and the covariates
    S.cols = me; S.nloci = 0; S.beta_len = 0; S.mu_computed = 0;
    S.ll = NAN;
    // p_proposal flat (src/sampler.hpp:351-362)
    p0[c] = ga / (ga + gb) * m_g / m_g;
    // vectors: xy[0] = sum y, d[0] = inv_tau2_e const
    std::fill(vec.begin(), vec.end(), 0.0);
    std::fill(idx.begin(), idx.end(), 0);
    // the fixed columns (src/model.hpp:92-105): E'y, E'E and the covariate
    // precisions, each dot summed in individual order; column 0 is the
    // constant (its dots are sums)
    const uint32_t nn_ = x->n;
    auto ecol = [&](int c, uint32_t i) { return c == 0 ? 1.0 : x->e_host[(size_t)(c - 1) * nn_ + i]; };
    for (int c = 0; c < me; ++c) {
      double sy = 0.0;
      for (uint32_t i = 0; i < nn_; ++i) sy += ecol(c, i) * x->y_host[i];
      vec[c] = sy;
      vec[2 * G + c] = c == 0 ? o.inv_tau2_e_const_val : o.inv_tau2_e_val;
      idx[3 * G + c] = c;     // slot
      idx[4 * G + c] = -1;    // no SNP
      idx[5 * G + c] = 4;     // covariate marker (src/model.hpp:98-101)
    }
    BMG_HIP(hipMemcpyAsync(A.vec, vec.data(), sizeof(double) * 5 * G, hipMemcpyHostToDevice, b->stream));
    BMG_HIP(hipMemcpyAsync(A.idx, idx.data(), sizeof(int32_t) * 8 * G, hipMemcpyHostToDevice, b->stream));
    std::vector<double> ee((size_t)me * me);
    for (int r = 0; r < me; ++r)
      for (int c = r; c < me; ++c) {
        double sum = 0.0;
        for (uint32_t i = 0; i < nn_; ++i) sum += ecol(r, i) * ecol(c, i);
        ee[(size_t)r * me + c] = ee[(size_t)c * me + r] = sum;
      }
    for (int r = 0; r < me; ++r)
      BMG_HIP(hipMemcpyAsync(A.xxs + (size_t)r * G, ee.data() + (size_t)r * me, sizeof(double) * me,
                             hipMemcpyHostToDevice, b->stream));
    std::fill(adapt.begin(), adapt.end(), 0.0);
    geometric_cdf_h((int)Kc.max_ms, o.p_move_size, adapt.data());
    geometric_cdf_h((int)(Kc.max_ms + 1) / 2, o.p_move_size_nbs, adapt.data() + 256);
    geometric_cdf_h((int)Kc.max_ms, o.p_move_size_nbc, adapt.data() + 512);
    BMG_HIP(hipMemcpyAsync(A.adapt, adapt.data(), sizeof(double) * 6 * 256, hipMemcpyHostToDevice, b->stream));
    BMG_HIP(hipStreamSynchronize(b->stream));
  }
  // online learned proposals: one learner per dataset that has chains
  b->learn.assign(n_ctx, nullptr);
  b->learn_on.assign(n_ctx, 0);
  if (b->cfg.proposal & ~3) {
    bmg_batch_destroy(b);
    set_error("proposal: 0 plain, 1 learned swap, 2 learned per-SNP proposal, 3 both");
    return BMG_E_ARG;
  }
  if (b->cfg.proposal != 0) {
    bmg_batch_cfg& bc = b->cfg;
    if (bc.learn_window <= 0) bc.learn_window = 20;
    if (bc.learn_steps <= 0) bc.learn_steps = 100;
    if (bc.learn_batch <= 0) bc.learn_batch = 4096;
    if (!(bc.learn_rate > 0)) bc.learn_rate = 1e-3;
    if (!(bc.p_learned > 0)) bc.p_learned = 0.9;
    if (bc.learn_window > BMG_LH_MAX || bc.p_learned > 1.0 || bc.learn_batch > (1 << 20)) {
      bmg_batch_destroy(b); set_error("learned proposal: window <= 32, p_learned <= 1, batch <= 2^20");
      return BMG_E_ARG;
    }
    for (int ds = 0; ds < n_ctx; ++ds) {
      int nc = 0;
      for (int c = 0; c < n_chains; ++c) nc += chain_ctx[c] == ds;
      if (!nc) continue;
      LearnCfg lc{bc.learn_window, bc.learn_steps, bc.learn_batch, (float)bc.learn_rate,
                  opts[0].seed * 7919u + (uint32_t)ds, bc.proposal & 1, (bc.proposal & 2) ? 1 : 0};
      int rc = learn_create(&b->learn[ds], ctxs[ds], b->h_data[ds], lc, nc, b->stream);
      if (rc) { bmg_batch_destroy(b); return rc; }
      b->h_data[ds].lw = learn_table(b->learn[ds]);
      b->h_data[ds].lH = bc.learn_window;
    }
  }
  BMG_HIP(hipMalloc(&b->d_data, sizeof(DevData) * n_ctx));
  BMG_HIP(hipMalloc(&b->d_consts, sizeof(ChainConst) * n_chains));
  BMG_HIP(hipMalloc(&b->d_scal, sizeof(ChainScal) * n_chains));
  BMG_HIP(hipMalloc(&b->d_arrays, sizeof(ChainArrays) * n_chains));
  BMG_HIP(hipMalloc(&b->rsum, sizeof(double) * 16));
  BMG_HIP(hipMalloc(&b->d_list, sizeof(int) * std::max(n_chains, 1024)));
  BMG_HIP(hipMemcpy(b->d_data, b->h_data.data(), sizeof(DevData) * n_ctx, hipMemcpyHostToDevice));
  BMG_HIP(hipMemcpy(b->d_consts, b->h_consts.data(), sizeof(ChainConst) * n_chains, hipMemcpyHostToDevice));
  BMG_HIP(hipMemcpy(b->d_scal, scal.data(), sizeof(ChainScal) * n_chains, hipMemcpyHostToDevice));
  BMG_HIP(hipMemcpy(b->d_arrays, b->h_arrays.data(), sizeof(ChainArrays) * n_chains, hipMemcpyHostToDevice));
  double* d_p0 = nullptr;
  BMG_HIP(hipMalloc(&d_p0, sizeof(double) * n_chains));
  BMG_HIP(hipMemcpy(d_p0, p0.data(), sizeof(double) * n_chains, hipMemcpyHostToDevice));
  uint64_t maxlen = 0;
  for (auto& D : b->h_data) maxlen = std::max(maxlen, D.lev_len[0]);
  hipLaunchKernelGGL(k_init_chain_arrays, dim3((unsigned)std::min<uint64_t>((maxlen + 255) / 256, 4096), n_chains),
                     dim3(256), 0, b->stream, b->d_data, b->d_consts, b->d_arrays, d_p0);
  BMG_HIP(hipGetLastError());
  // build the upper tree levels of every chain, dataset by dataset
  for (size_t ds = 0; ds < b->h_data.size(); ++ds) {
    std::vector<int> list;
    for (int c = 0; c < n_chains; ++c) if (b->chain_ds[c] == (int)ds) list.push_back(c);
    for (size_t s0 = 0; s0 < list.size(); s0 += 1024) {
      int nv = (int)std::min<size_t>(1024, list.size() - s0);
      BMG_HIP(hipMemcpyAsync(b->d_list, list.data() + s0, sizeof(int) * nv, hipMemcpyHostToDevice, b->stream));
      if (launch_tree_build(b->d_data, b->d_consts, b->d_scal, b->d_arrays, b->d_list, nv,
                            b->h_data[ds].lev_len, b->h_data[ds].top, 0, 0, b->stream)) {
        bmg_batch_destroy(b); set_error("tree build failed"); return BMG_E_HIP;
      }
      BMG_HIP(hipStreamSynchronize(b->stream));
    }
  }
  hipFree(d_p0);
  *out = b;
  return BMG_OK;
}

extern "C" int bmg_batch_create(bmg_ctx** ctxs, int n_ctx, int n_chains, const int32_t* chain_ctx,
                                const bmg_opts* opts, const bmg_batch_cfg* cfg, bmg_batch** out) {
  return batch_create(ctxs, n_ctx, n_chains, chain_ctx, opts, cfg, 0, out);
}

extern "C" int bmg_batch_setup(bmg_batch* b) {
  if (!b) { set_error("null batch"); return BMG_E_ARG; }
  if (b->setup_done) { set_error("setup already done"); return BMG_E_STATE; }
  hipSetDevice(b->device);
  // x_j . y of every SNP for each coding the dataset's chains add with (y is
  // fixed, so each add reads one double instead of the whole of y)
  for (size_t ds = 0; ds < b->h_data.size(); ++ds) {
    DevData& D = b->h_data[ds];
    const int codings[2] = {D.type == BMG_AH ? 0 : D.type, D.type == BMG_AH ? 1 : -1};
    for (int cd : codings) {
      if (cd < 0 || D.xy[cd]) continue;
      double* t = nullptr;
      BMG_HIP(hipMalloc(&t, sizeof(double) * D.p));
      b->xy_dev.push_back(t);
      if (launch_xy_all(D, cd, t, b->stream)) { set_error("x.y table launch failed"); return BMG_E_HIP; }
      D.xy[cd] = t;
    }
  }
  BMG_HIP(hipMemcpyAsync(b->d_data, b->h_data.data(), sizeof(DevData) * b->h_data.size(), hipMemcpyHostToDevice,
                         b->stream));
  int rc = launch_chain_kernel(b, 0, 1);
  if (rc) return rc;
  BMG_HIP(hipStreamSynchronize(b->stream));
  b->setup_done = 1;
  return BMG_OK;
}

// --- steady-state graph ---------------------------------------------------
// After burn-in (every chain's n_rao_burnin at 0, every learner frozen) each
// RB point triggers the same device work for every chain: a scan of every
// chain of every dataset, the element-wise update and the tree rebuild.  A
// segment -- the chain kernel over n_rao iterations plus that RB block -- is
// then captured once into a hipGraph and replayed; the host decides nothing
// between segments.  The captured sequence is the eager one with the chain
// lists read from a device array and without host round trips.
static int steady_state(bmg_batch* b) {
  std::vector<ChainScal> cs(b->n_chains);
  BMG_HIP(hipStreamSynchronize(b->stream));
  BMG_HIP(hipMemcpy(cs.data(), b->d_scal, sizeof(ChainScal) * b->n_chains, hipMemcpyDeviceToHost));
  // (capacity and tree flags only record rejected proposals; a stopped
  // chain or a full trace buffer ends the graph path, as it ends the run)
  for (const ChainScal& c : cs)
    if (c.n_rao_burnin > 0 || c.rb_pending || (c.err & (ERR_TRACE | ERR_STOP))) return 0;
  for (LearnDS* L : b->learn)
    if (L && !learn_frozen(L)) return 0;
  return 1;
}

// the buffers a captured segment points at; a change means a rebuild
static std::vector<const void*> graph_signature(bmg_batch* b) {
  std::vector<const void*> v = {b->r, b->xa, b->xh, b->rsum, b->d_glist};
  for (bmg_ctx* c : b->ctxs) { v.push_back(c->dig); v.push_back(c->acc); v.push_back(c->scale); }
  return v;
}

static int build_graph(bmg_batch* b) {
  if (b->graph_exec) { hipGraphExecDestroy(b->graph_exec); b->graph_exec = nullptr; }
  if (b->graph) { hipGraphDestroy(b->graph); b->graph = nullptr; }
  const int digits = b->cfg.scan_digits > 0 ? b->cfg.scan_digits : 5;
  const int per = 16 / digits;
  // chain lists and scratch at their final sizes before the capture
  std::vector<int> order;
  std::vector<size_t> ds_off;
  for (size_t ds = 0; ds < b->ctxs.size(); ++ds) {
    ds_off.push_back(order.size());
    for (int c = 0; c < b->n_chains; ++c) if (b->chain_ds[c] == (int)ds) order.push_back(c);
  }
  ds_off.push_back(order.size());
  if (!b->d_glist) BMG_HIP(hipMalloc(&b->d_glist, sizeof(int) * b->n_chains));
  BMG_HIP(hipMemcpy(b->d_glist, order.data(), sizeof(int) * order.size(), hipMemcpyHostToDevice));
  for (size_t ds = 0; ds < b->ctxs.size(); ++ds)
    if (ensure_rb_scratch(b, per, b->h_data[ds].n_pad, b->h_data[ds].p)) return BMG_E_HIP;
  int scans = 0;
  BMG_HIP(hipStreamBeginCapture(b->stream, hipStreamCaptureModeRelaxed));
  int rc = launch_chain_kernel(b, b->n_rao, 0);
  for (size_t ds = 0; ds < b->ctxs.size() && !rc; ++ds) {
    bmg_ctx* c = b->ctxs[ds];
    const DevData& D = b->h_data[ds];
    c->no_tune = 1;
    for (size_t s0 = ds_off[ds]; s0 < ds_off[ds + 1] && !rc; s0 += per) {
      const int nv = (int)std::min<size_t>(per, ds_off[ds + 1] - s0);
      const int* list = b->d_glist + s0;
      if (launch_residuals(c->y, D.n, D.n_pad, b->d_arrays, list, nv, b->r, b->rsum, b->stream)) {
        set_error("residual kernel launch failed"); rc = BMG_E_HIP; break;
      }
      const int t = D.type;
      rc = launch_scan(c, b->r, nv, t == BMG_AH ? 0 : t, digits, b->xa, b->stream);
      if (!rc && t == BMG_AH) rc = launch_scan(c, b->r, nv, 1, digits, b->xh, b->stream);
      scans += (t == BMG_AH) ? 2 : 1;
      if (!rc && launch_rb_finalize(b->d_data, b->d_consts, b->d_scal, b->d_arrays, list, nv, BMG_KMAXG, D.p,
                                    b->xa, b->xh, b->rsum, b->stream)) {
        set_error("rb finalize launch failed"); rc = BMG_E_HIP;
      }
    }
    for (size_t s0 = ds_off[ds]; s0 < ds_off[ds + 1] && !rc; s0 += 1024) {
      const int nv = (int)std::min<size_t>(1024, ds_off[ds + 1] - s0);
      const int* list = b->d_glist + s0;
      if (launch_rb_update(b->d_data, b->d_consts, b->d_scal, b->d_arrays, list, nv, D.p, b->stream) ||
          launch_tree_build(b->d_data, b->d_consts, b->d_scal, b->d_arrays, list, nv, D.lev_len, D.top, 1, 1,
                            b->stream)) {
        set_error("rb update launch failed"); rc = BMG_E_HIP;
      }
    }
    c->no_tune = 0;
  }
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(b->stream, &g);
  if (rc) { if (g) hipGraphDestroy(g); return rc; }
  if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
  b->graph = g;
  e = hipGraphInstantiate(&b->graph_exec, g, nullptr, nullptr, 0);
  if (e != hipSuccess) return hip_fail(e, "hipGraphInstantiate");
  b->graph_sig = graph_signature(b);
  b->graph_scans = scans;
  b->n_builds += 1;
  return BMG_OK;
}

// replays `segments` captured segments; the batch is at an RB boundary in
// its steady state
static int run_graph(bmg_batch* b, uint64_t segments) {
  if (!b->graph_exec || graph_signature(b) != b->graph_sig) {
    int rc = build_graph(b);
    if (rc) return rc;
  }
  hipEvent_t g0, g1;
  BMG_HIP(hipEventCreate(&g0));
  BMG_HIP(hipEventCreate(&g1));
  BMG_HIP(hipEventRecord(g0, b->stream));
  for (uint64_t k = 0; k < segments; ++k) BMG_HIP(hipGraphLaunch(b->graph_exec, b->stream));
  BMG_HIP(hipEventRecord(g1, b->stream));
  BMG_HIP(hipEventSynchronize(g1));
  float ms = 0;
  (void)hipEventElapsedTime(&ms, g0, g1);
  (void)hipEventDestroy(g0);
  (void)hipEventDestroy(g1);
  b->t_graph += ms;
  b->n_replays += (double)segments;
  b->n_scans += (int)(segments * b->graph_scans);
  b->iter += segments * b->n_rao;
  return BMG_OK;
}

extern "C" int bmg_batch_run(bmg_batch* b, uint64_t n_iter) {
  if (!b) { set_error("null batch"); return BMG_E_ARG; }
  if (!b->setup_done) { set_error("call bmg_batch_setup first"); return BMG_E_STATE; }
  hipSetDevice(b->device);
  hipEvent_t e0, e1, c0, c1;
  hipEventCreate(&e0); hipEventCreate(&e1); hipEventCreate(&c0); hipEventCreate(&c1);
  hipEventRecord(e0, b->stream);
  uint64_t remaining = n_iter;
  int rc = BMG_OK;
  while (remaining > 0) {
    if (b->cfg.use_graph && b->iter % b->n_rao == 0 && remaining >= b->n_rao && steady_state(b)) {
      const uint64_t segs = remaining / b->n_rao;
      rc = run_graph(b, segs);
      if (rc) break;
      remaining -= segs * b->n_rao;
      continue;
    }
    uint64_t next_rb = (b->iter / b->n_rao + 1) * b->n_rao;
    uint64_t run = std::min<uint64_t>(remaining, next_rb - b->iter);
    hipEventRecord(c0, b->stream);
    rc = launch_chain_kernel(b, run, 0);
    if (rc) break;
    hipEventRecord(c1, b->stream);
    hipEventSynchronize(c1);
    float ms = 0; hipEventElapsedTime(&ms, c0, c1); b->t_chain += ms;
    b->iter += run;
    remaining -= run;
    if (b->iter % b->n_rao == 0) {
      hipEventRecord(c0, b->stream);
      rc = rb_block(b);
      if (rc) break;
      hipEventRecord(c1, b->stream);
      hipEventSynchronize(c1);
      ms = 0; hipEventElapsedTime(&ms, c0, c1); b->t_scan += ms;
    }
  }
  hipEventRecord(e1, b->stream);
  hipEventSynchronize(e1);
  float tt = 0; hipEventElapsedTime(&tt, e0, e1); b->t_total += tt;
  hipEventDestroy(e0); hipEventDestroy(e1); hipEventDestroy(c0); hipEventDestroy(c1);
  if (rc) return rc;
  hipError_t e = hipStreamSynchronize(b->stream);
  if (e != hipSuccess) return hip_fail(e, "bmg_batch_run");
  std::vector<ChainScal> cs(b->n_chains);
  BMG_HIP(hipMemcpy(cs.data(), b->d_scal, sizeof(ChainScal) * b->n_chains, hipMemcpyDeviceToHost));
  return numeric_error(cs);
}

extern "C" int bmg_batch_graph_stats(bmg_batch* b, double out[3]) {
  if (!b || !out) { set_error("null argument"); return BMG_E_ARG; }
  out[0] = b->n_replays; out[1] = b->t_graph; out[2] = b->n_builds;
  return BMG_OK;
}

// learned swap on/off for the chains of dataset ds (see move_learned_swap)
static int set_learn_on(bmg_batch* b, int ds, int on, double p_learned) {
  const int32_t v = on;
  for (int c = 0; c < b->n_chains; ++c) {
    if (b->chain_ds[c] != ds) continue;
    b->h_consts[c].p_learned = p_learned;
    BMG_HIP(hipMemcpyAsync((char*)(b->d_scal + c) + offsetof(ChainScal, learn_on), &v, sizeof(v),
                           hipMemcpyHostToDevice, b->stream));
  }
  BMG_HIP(hipMemcpyAsync(b->d_consts, b->h_consts.data(), sizeof(ChainConst) * b->n_chains,
                         hipMemcpyHostToDevice, b->stream));
  BMG_HIP(hipStreamSynchronize(b->stream));
  return BMG_OK;
}

extern "C" int bmg_batch_set_learned(bmg_batch* b, int ds, const float* w, int H, double p_learned) {
  if (!b || ds < 0 || ds >= (int)b->ctxs.size()) { set_error("bad batch or dataset"); return BMG_E_ARG; }
  if (H < 1 || H > BMG_LH_MAX || !(p_learned >= 0.0 && p_learned <= 1.0)) {
    set_error("learned swap: 1 <= H <= 32 and 0 <= p_learned <= 1"); return BMG_E_ARG;
  }
  if (ds < (int)b->learn.size() && b->learn[ds]) {
    set_error("this dataset's proposals are learned online (proposal != 0)"); return BMG_E_STATE;
  }
  BMG_HIP(hipSetDevice(b->device));
  BMG_HIP(hipStreamSynchronize(b->stream));
  if (!w) return set_learn_on(b, ds, 0, 0.0);
  if (b->lw_dev.size() < b->ctxs.size()) b->lw_dev.resize(b->ctxs.size(), nullptr);
  DevData& D = b->h_data[ds];
  size_t cnt = (size_t)D.p * 2 * H;
  if (b->lw_dev[ds] && D.lH != H) { hipFree(b->lw_dev[ds]); b->lw_dev[ds] = nullptr; }
  if (!b->lw_dev[ds]) BMG_HIP(hipMalloc(&b->lw_dev[ds], sizeof(float) * cnt));
  std::vector<float> hw(w, w + cnt);
  for (float& x : hw) x = x > 0.0f && std::isfinite(x) ? x : 0.0f;
  BMG_HIP(hipMemcpy(b->lw_dev[ds], hw.data(), sizeof(float) * cnt, hipMemcpyHostToDevice));
  D.lw = b->lw_dev[ds]; D.lH = H;
  BMG_HIP(hipMemcpy(b->d_data, b->h_data.data(), sizeof(DevData) * b->h_data.size(), hipMemcpyHostToDevice));
  return set_learn_on(b, ds, 1, p_learned);
}

static LearnDS* learner(bmg_batch* b, int ds) {
  if (!b || ds < 0 || ds >= (int)b->learn.size() || !b->learn[ds]) {
    set_error("no learned proposal for this dataset (proposal != 0 at create)");
    return nullptr;
  }
  hipSetDevice(b->device);
  return b->learn[ds];
}

extern "C" int bmg_batch_learn_state(bmg_batch* b, int ds, double out[4]) {
  LearnDS* L = learner(b, ds);
  if (!L || !out) return BMG_E_ARG;
  int64_t cnt = 0;
  int rc = learn_samples(L, nullptr, nullptr, nullptr, 0, &cnt, b->stream);
  if (rc) return rc;
  out[0] = (double)learn_rounds(L); out[1] = learn_frozen(L); out[2] = (double)cnt; out[3] = learn_H(L);
  return BMG_OK;
}

extern "C" int bmg_batch_learn_table(bmg_batch* b, int ds, float* w) {
  LearnDS* L = learner(b, ds);
  if (!L || !w) return BMG_E_ARG;
  BMG_HIP(hipStreamSynchronize(b->stream));
  BMG_HIP(hipMemcpy(w, learn_table(L), sizeof(float) * b->h_data[ds].p * 2 * learn_H(L), hipMemcpyDeviceToHost));
  return BMG_OK;
}

extern "C" int bmg_batch_learn_samples(bmg_batch* b, int ds, uint32_t* j, int32_t* off, float* label,
                                       int64_t cap, int64_t* count) {
  LearnDS* L = learner(b, ds);
  if (!L || !count) return BMG_E_ARG;
  return learn_samples(L, j, off, label, cap, count, b->stream);
}

extern "C" int bmg_batch_learn_params(bmg_batch* b, int ds, float* params) {
  LearnDS* L = learner(b, ds);
  if (!L || !params) return BMG_E_ARG;
  return learn_params(L, params, b->stream);
}

extern "C" int bmg_batch_learn_static(bmg_batch* b, int ds, float* z, float* pbar, int32_t* band) {
  LearnDS* L = learner(b, ds);
  if (!L) return BMG_E_ARG;
  return learn_static(L, z, pbar, band, b->stream);
}

extern "C" int bmg_batch_learn_snp_features(bmg_batch* b, int ds, int chain, const uint32_t* j, int64_t rows,
                                            float* out) {
  LearnDS* L = learner(b, ds);
  if (!L) return BMG_E_ARG;
  if (chain < 0 || chain >= b->n_chains || b->chain_ds[chain] != ds || !j || !out || rows < 1) {
    set_error("bad argument"); return BMG_E_ARG;
  }
  const uint64_t p = b->h_data[ds].p;
  for (int64_t i = 0; i < rows; ++i)
    if (j[i] >= p) { set_error("SNP index out of range"); return BMG_E_ARG; }
  BMG_HIP(hipStreamSynchronize(b->stream));
  return learn_snp_features(L, b->d_arrays, chain, j, rows, out, b->stream);
}

extern "C" int bmg_batch_learn_features(bmg_batch* b, int ds, const uint32_t* j, const int32_t* off,
                                        int64_t rows, float* out) {
  LearnDS* L = learner(b, ds);
  if (!L || !j || !off || !out || rows < 1) return BMG_E_ARG;
  const uint64_t p = b->h_data[ds].p;
  const int H = learn_H(L);
  for (int64_t i = 0; i < rows; ++i) {
    int64_t k = (int64_t)j[i] + off[i];
    if (j[i] >= p || off[i] == 0 || off[i] < -H || off[i] > H || k < 0 || k >= (int64_t)p) {
      set_error("feature rows need 0 < |offset| <= H with both SNPs inside the genome"); return BMG_E_ARG;
    }
  }
  return learn_features(L, j, off, rows, out, b->stream);
}

extern "C" int bmg_batch_profile(bmg_batch* b, int chain, double out[24]) {
  if (!b || chain < 0 || chain >= b->n_chains || !out) { set_error("bad argument"); return BMG_E_ARG; }
  BMG_HIP(hipSetDevice(b->device));
  BMG_HIP(hipStreamSynchronize(b->stream));
  ChainScal cs;
  BMG_HIP(hipMemcpy(&cs, b->d_scal + chain, sizeof(cs), hipMemcpyDeviceToHost));
  int khz = 0;
  BMG_HIP(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, b->device));
  for (int q = 0; q < 24; ++q)
    out[q] = (q == 12 || q == 13) ? (double)cs.prof[q] : (khz > 0 ? (double)cs.prof[q] / (double)khz : 0.0);
  return BMG_OK;
}

extern "C" int bmg_batch_sync(bmg_batch* b) {
  if (!b) { set_error("null batch"); return BMG_E_ARG; }
  BMG_HIP(hipStreamSynchronize(b->stream));
  return BMG_OK;
}

extern "C" int bmg_batch_timing(bmg_batch* b, double out[4]) {
  if (!b || !out) { set_error("null argument"); return BMG_E_ARG; }
  out[0] = b->t_total; out[1] = b->t_chain; out[2] = b->t_scan; out[3] = b->n_scans;
  return BMG_OK;
}

static int get_scal(bmg_batch* b, int chain, ChainScal* s) {
  if (!b || chain < 0 || chain >= b->n_chains) { set_error("bad chain"); return BMG_E_ARG; }
  BMG_HIP(hipStreamSynchronize(b->stream));
  BMG_HIP(hipMemcpy(s, b->d_scal + chain, sizeof(ChainScal), hipMemcpyDeviceToHost));
  return BMG_OK;
}

extern "C" int bmg_batch_state(bmg_batch* b, int chain, double out[8]) {
  ChainScal s;
  int rc = get_scal(b, chain, &s);
  if (rc) return rc;
  out[0] = s.ll; out[1] = s.nloci; out[2] = s.sigma2; out[3] = s.alpha; out[4] = s.p_ms;
  out[5] = (double)s.n_accepted; out[6] = (double)b->iter; out[7] = (double)s.th_loci_len;
  return BMG_OK;
}

extern "C" int bmg_batch_loci(bmg_batch* b, int chain, uint32_t* loci, int cap, int* k) {
  ChainScal s;
  int rc = get_scal(b, chain, &s);
  if (rc) return rc;
  *k = s.nloci;
  std::vector<int32_t> buf(s.nloci);
  if (s.nloci) BMG_HIP(hipMemcpy(buf.data(), b->h_arrays[chain].idx, sizeof(int32_t) * s.nloci, hipMemcpyDeviceToHost));
  for (int i = 0; i < s.nloci && i < cap; ++i) loci[i] = (uint32_t)buf[i];
  return BMG_OK;
}

static int copy_p(bmg_batch* b, int chain, const double* src, double* out) {
  if (!b || chain < 0 || chain >= b->n_chains || !out) { set_error("bad argument"); return BMG_E_ARG; }
  BMG_HIP(hipStreamSynchronize(b->stream));
  BMG_HIP(hipMemcpy(out, src, sizeof(double) * b->h_data[b->chain_ds[chain]].p, hipMemcpyDeviceToHost));
  return BMG_OK;
}

extern "C" int bmg_batch_p_rao(bmg_batch* b, int chain, double* out) {
  if (!b || chain < 0 || chain >= b->n_chains) { set_error("bad chain"); return BMG_E_ARG; }
  return copy_p(b, chain, b->h_arrays[chain].p_rao, out);
}

extern "C" int bmg_batch_p_proposal(bmg_batch* b, int chain, double* out) {
  if (!b || chain < 0 || chain >= b->n_chains) { set_error("bad chain"); return BMG_E_ARG; }
  return copy_p(b, chain, b->h_arrays[chain].p_proposal, out);
}

extern "C" int bmg_batch_set_p_proposal(bmg_batch* b, int chain, const double* p) {
  if (!b || chain < 0 || chain >= b->n_chains || !p) { set_error("bad argument"); return BMG_E_ARG; }
  if (!b->setup_done) { set_error("call bmg_batch_setup first"); return BMG_E_STATE; }
  const int ds = b->chain_ds[chain];
  const uint64_t np = b->h_data[ds].p;
  for (uint64_t j = 0; j < np; ++j)
    if (!(p[j] >= 0.0 && p[j] <= 1.0)) { set_error("proposal table: values must lie in [0, 1]"); return BMG_E_ARG; }
  hipSetDevice(b->device);
  BMG_HIP(hipStreamSynchronize(b->stream));
  double* d_p = nullptr;
  BMG_HIP(hipMalloc(&d_p, sizeof(double) * np));
  int rc = BMG_OK;
  if (hipMemcpy(d_p, p, sizeof(double) * np, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(b->d_list, &chain, sizeof(int), hipMemcpyHostToDevice) != hipSuccess) {
    set_error("proposal table: copy failed"); rc = BMG_E_HIP;
  }
  if (!rc) {
    if (launch_set_proposal(b->d_data, b->d_consts, b->d_arrays, chain, d_p, np, b->stream) ||
        launch_tree_build(b->d_data, b->d_consts, b->d_scal, b->d_arrays, b->d_list, 1, b->h_data[ds].lev_len,
                          b->h_data[ds].top, 0, 0, b->stream) ||
        hipStreamSynchronize(b->stream) != hipSuccess) {
      set_error("proposal table: device update failed"); rc = BMG_E_HIP;
    }
  }
  hipFree(d_p);
  return rc;
}

extern "C" int bmg_batch_trace(bmg_batch* b, int chain, uint8_t* move, uint8_t* jd, uint8_t* ms,
                               uint32_t* size, double* loglik, int64_t cap, int64_t* len) {
  ChainScal s;
  int rc = get_scal(b, chain, &s);
  if (rc) return rc;
  const ChainArrays& A = b->h_arrays[chain];
  int64_t n = (int64_t)std::min<uint64_t>(s.tr_count, b->h_consts[chain].tr_cap);
  if (n > cap) n = cap;
  *len = n;
  if (n > 0) {
    if (move) BMG_HIP(hipMemcpy(move, A.tr_move, n, hipMemcpyDeviceToHost));
    if (jd) BMG_HIP(hipMemcpy(jd, A.tr_jd, n, hipMemcpyDeviceToHost));
    if (ms) BMG_HIP(hipMemcpy(ms, A.tr_ms, n, hipMemcpyDeviceToHost));
    if (size) BMG_HIP(hipMemcpy(size, A.tr_size, 4 * n, hipMemcpyDeviceToHost));
    if (loglik) BMG_HIP(hipMemcpy(loglik, A.tr_ll, 8 * n, hipMemcpyDeviceToHost));
  }
  return BMG_OK;
}

extern "C" int bmg_batch_thinned(bmg_batch* b, int chain, uint32_t* sizes, int64_t cap_sizes, uint32_t* loci,
                                 int64_t cap_loci, double* loglik, double* sigma2, double* alpha,
                                 int64_t* n_samples, int64_t* n_loci) {
  ChainScal s;
  int rc = get_scal(b, chain, &s);
  if (rc) return rc;
  const ChainArrays& A = b->h_arrays[chain];
  int64_t ns = (int64_t)s.th_count, nl = (int64_t)s.th_loci_len;
  *n_samples = ns; *n_loci = nl;
  int64_t cs = std::min(ns, cap_sizes), cl = std::min(nl, cap_loci);
  if (cs > 0) {
    if (sizes) BMG_HIP(hipMemcpy(sizes, A.th_size, 4 * cs, hipMemcpyDeviceToHost));
    if (loglik) BMG_HIP(hipMemcpy(loglik, A.th_ll, 8 * cs, hipMemcpyDeviceToHost));
    if (sigma2) BMG_HIP(hipMemcpy(sigma2, A.th_s2, 8 * cs, hipMemcpyDeviceToHost));
    if (alpha) BMG_HIP(hipMemcpy(alpha, A.th_alpha, 8 * cs, hipMemcpyDeviceToHost));
  }
  if (cl > 0 && loci) BMG_HIP(hipMemcpy(loci, A.th_loci, 4 * cl, hipMemcpyDeviceToHost));
  return BMG_OK;
}

// --- output files of Sampler::sample (src/sampler.cpp:560-584,687-728,846-858)
static int write_file(const std::string& path, const void* data, size_t bytes) {
  FILE* f = fopen(path.c_str(), "wb");
  if (!f) { set_error("cannot open " + path); return BMG_E_ARG; }
  size_t w = bytes ? fwrite(data, 1, bytes, f) : 0;
  int bad = fclose(f) != 0 || w != bytes;
  if (bad) { set_error("cannot write " + path); return BMG_E_ARG; }
  return BMG_OK;
}

extern "C" int bmg_batch_write_outputs(bmg_batch* b, int chain, const char* basename) {
  if (!basename) { set_error("null basename"); return BMG_E_ARG; }
  ChainScal s;
  int rc = get_scal(b, chain, &s);
  if (rc) return rc;
  if (!b->cfg.trace_every_iter) {
    set_error("the per-iteration files need trace_every_iter = 1"); return BMG_E_STATE;
  }
  if (s.err & ERR_TRACE) { set_error("trace buffers overflowed: the files would be incomplete"); return BMG_E_CAPACITY; }
  const ChainArrays& A = b->h_arrays[chain];
  const ChainConst& K = b->h_consts[chain];
  const std::string base(basename);
  const size_t ns = (size_t)s.th_count, nl = (size_t)s.th_loci_len, nt = (size_t)s.tr_count;
  std::vector<uint32_t> sizes(ns), loci(nl);
  std::vector<double> ll(ns), s2(ns), al(ns), pve(3 * ns), lp(ns);
  std::vector<uint8_t> mv(nt), jd(nt), ms(nt);
  std::vector<double> rao(b->h_data[b->chain_ds[chain]].p);
  if (ns) {
    BMG_HIP(hipMemcpy(sizes.data(), A.th_size, 4 * ns, hipMemcpyDeviceToHost));
    BMG_HIP(hipMemcpy(ll.data(), A.th_ll, 8 * ns, hipMemcpyDeviceToHost));
    BMG_HIP(hipMemcpy(s2.data(), A.th_s2, 8 * ns, hipMemcpyDeviceToHost));
    BMG_HIP(hipMemcpy(al.data(), A.th_alpha, 8 * ns, hipMemcpyDeviceToHost));
    BMG_HIP(hipMemcpy(pve.data(), A.th_pve, 8 * 3 * ns, hipMemcpyDeviceToHost));
  }
  if (nl) BMG_HIP(hipMemcpy(loci.data(), A.th_loci, 4 * nl, hipMemcpyDeviceToHost));
  if (nt) {
    BMG_HIP(hipMemcpy(mv.data(), A.tr_move, nt, hipMemcpyDeviceToHost));
    BMG_HIP(hipMemcpy(jd.data(), A.tr_jd, nt, hipMemcpyDeviceToHost));
    BMG_HIP(hipMemcpy(ms.data(), A.tr_ms, nt, hipMemcpyDeviceToHost));
  }
  BMG_HIP(hipMemcpy(rao.data(), A.p_rao, 8 * rao.size(), hipMemcpyDeviceToHost));
  // the learned swap has no move code in the reference (3 is its type switch)
  for (uint8_t& m : mv) if (m == 3) m = BMG_MOVE_LEARNED;
  // log model prior (src/prior.hpp:169-183) of one effect type with n_g terms:
  // lgamma(tp + n) + lgamma(a + n) + lgamma(b + m - n) - lgamma(tp_sum + n)
  for (size_t t = 0; t < ns; ++t) {
    const double n = (double)sizes[t];
    lp[t] = std::lgamma(K.tp_type + n) + (std::lgamma(K.g_a + n) + std::lgamma(K.g_b + K.m_g - n) -
                                          std::lgamma(K.tp_sum + n));
  }
  struct F { const char* name; const void* p; size_t bytes; };
  const F files[] = {
      {"_loci.dat", loci.data(), 4 * nl}, {"_modelsize.dat", sizes.data(), 4 * ns},
      {"_jumpdistance.dat", jd.data(), nt}, {"_log_likelihood.dat", ll.data(), 8 * ns},
      {"_log_prior.dat", lp.data(), 8 * ns}, {"_move_type.dat", mv.data(), nt},
      {"_move_size.dat", ms.data(), nt}, {"_pve.dat", pve.data(), 8 * 3 * ns},
      {"_alpha.dat", al.data(), 8 * ns}, {"_sigma2.dat", s2.data(), 8 * ns},
      {"_rao.dat", rao.data(), 8 * rao.size()}};
  for (const F& f : files) {
    rc = write_file(base + f.name, f.p, f.bytes);
    if (rc) return rc;
  }
  if (K.th_beta_cap) {
    // src/sampler.cpp:710-713: beta of every thinned sample, concatenated
    std::vector<double> beta(s.th_beta_len);
    if (!beta.empty())
      BMG_HIP(hipMemcpy(beta.data(), A.th_beta, 8 * beta.size(), hipMemcpyDeviceToHost));
    rc = write_file(base + "_beta.dat", beta.data(), 8 * beta.size());
    if (rc) return rc;
  }
  return BMG_OK;
}

extern "C" int bmg_batch_thinned_beta(bmg_batch* b, int chain, double* out, int64_t cap, int64_t* len) {
  ChainScal s;
  int rc = get_scal(b, chain, &s);
  if (rc) return rc;
  if (!b->h_consts[chain].th_beta_cap) { set_error("the chain does not save beta (save_beta = 0)"); return BMG_E_STATE; }
  if (len) *len = (int64_t)s.th_beta_len;
  const int64_t k = std::min<int64_t>(cap, (int64_t)s.th_beta_len);
  if (out && k > 0)
    BMG_HIP(hipMemcpy(out, b->h_arrays[chain].th_beta, 8 * k, hipMemcpyDeviceToHost));
  return BMG_OK;
}

extern "C" int bmg_batch_rb(bmg_batch* b, int chain, const double* y_hat, double* p_r) {
  ChainScal s;
  int rc = get_scal(b, chain, &s);
  if (rc) return rc;
  int ds = b->chain_ds[chain];
  const DevData& D = b->h_data[ds];
  BMG_HIP(hipMemset(b->h_arrays[chain].y_hat, 0, sizeof(double) * D.n_pad));
  BMG_HIP(hipMemcpy(b->h_arrays[chain].y_hat, y_hat, sizeof(double) * D.n, hipMemcpyHostToDevice));
  s.rb_call_idx = s.rb_calls;
  s.rb_calls++;
  BMG_HIP(hipMemcpy(b->d_scal + chain, &s, sizeof(ChainScal), hipMemcpyHostToDevice));
  std::vector<int> list{chain};
  rc = rb_block_ds(b, ds, list, 0);
  if (rc) return rc;
  BMG_HIP(hipStreamSynchronize(b->stream));
  BMG_HIP(hipMemcpy(p_r, b->h_arrays[chain].p_r, sizeof(double) * D.p, hipMemcpyDeviceToHost));
  return BMG_OK;
}

// ----------------------------------------------------------------- model --
extern "C" int bmg_model_kmax(void) { return 160; }

static int launch_model_op(bmg_batch* b, ModelOpArgs a) {
  a.data = b->d_data; a.consts = b->d_consts; a.scal = b->d_scal; a.arrays = b->d_arrays;
  switch (b->kmax) {
    case 64: hipLaunchKernelGGL(k_model_op<64>, dim3(1), dim3(256), 0, b->stream, a); break;
    case 128: hipLaunchKernelGGL(k_model_op<128>, dim3(1), dim3(256), 0, b->stream, a); break;
    case 160: hipLaunchKernelGGL(k_model_op<160>, dim3(1), dim3(256), 0, b->stream, a); break;
    default: hipLaunchKernelGGL(k_model_op<256>, dim3(1), dim3(256), 0, b->stream, a); break;
  }
  BMG_HIP(hipGetLastError());
  BMG_HIP(hipStreamSynchronize(b->stream));
  return BMG_OK;
}

extern "C" int bmg_model_create(bmg_ctx* ctx, const bmg_model_prior* prior, bmg_model** out) {
  if (!ctx || !prior || !out) { set_error("null argument"); return BMG_E_ARG; }
  if (ctx->n_types < 1) { set_error("dataset has no effect type"); return BMG_E_ARG; }
  if (ctx->m_e > 0) {
    set_error("a dataset with covariates needs their precisions: use bmg_model_create_sampler");
    return BMG_E_ARG;
  }
  // a model is a one-chain batch whose constants carry the given prior
  bmg_opts o{};
  o.n = ctx->n; o.m_g = ctx->p; o.do_n_iter = 1; o.n_rao = 1; o.thin = 1; o.n_sample_tau2_and_missing = 1;
  o.max_move_size = 1; o.max_SNP_neighborhood_size = 1; o.seed = 1;
  o.p_move_size = 0.5; o.p_move_size_nbs = 0.5; o.p_move_size_nbc = 0.5;
  o.n_types = 1; o.types[0] = ctx->types[0];
  o.e_qg = 1; o.var_qg = 1; o.nu_sigma2 = 1; o.s2_sigma2 = 1; o.mu_alpha = 1;
  for (int t = 0; t < 4; ++t) { o.nu_tau2[t] = 4; o.s2_tau2[t] = 1; }
  for (int t = 0; t < 5; ++t) o.types_prior[t] = 1;
  o.inv_tau2_e_const_val = prior->inv_tau2_e0;
  // the model API tolerates missing genotypes (they read as 0, the value
  // the reference's models see before missing values are sampled) and
  // terms of every type the dataset allows
  int32_t zero = 0;
  bmg_batch_cfg cfg{};
  bmg_batch* b = nullptr;
  int rc = batch_create(&ctx, 1, 1, &zero, &o, &cfg, 1, &b);
  if (rc) return rc;
  ChainConst& K = b->h_consts[0];
  K.n_plus_nu = prior->n_plus_nu;
  K.minus_n_plus_nu_2 = prior->minus_n_plus_nu_2;
  K.nus2_plus_yy = prior->nus2_plus_yy;
  K.nus2 = prior->nus2_plus_yy - ctx->yy;
  BMG_HIP(hipMemcpy(b->d_consts, &K, sizeof(ChainConst), hipMemcpyHostToDevice));
  bmg_model* m = new bmg_model();
  m->b = b;
  ModelOpArgs a{};
  a.op = 2;
  rc = launch_model_op(b, a);
  if (rc) { bmg_batch_destroy(b); delete m; return rc; }
  *out = m;
  return BMG_OK;
}

extern "C" int bmg_model_create_sampler(bmg_ctx* ctx, const bmg_opts* opts, bmg_model** out) {
  if (!ctx || !opts || !out) { set_error("null argument"); return BMG_E_ARG; }
  if (ctx->n_types < 1) { set_error("dataset has no effect type"); return BMG_E_ARG; }
  // the model of one Sampler: its prior and its random stream (key = seed,
  // chain) from the options, as a batch chain before set-up
  bmg_opts o = *opts;
  o.do_n_iter = 1;
  int32_t zero = 0;
  bmg_batch_cfg cfg{};
  bmg_batch* b = nullptr;
  int rc = batch_create(&ctx, 1, 1, &zero, &o, &cfg, 1, &b);
  if (rc) return rc;
  bmg_model* m = new bmg_model();
  m->b = b;
  ModelOpArgs a{};
  a.op = 2;
  rc = launch_model_op(b, a);
  if (rc) { bmg_batch_destroy(b); delete m; return rc; }
  *out = m;
  return BMG_OK;
}

extern "C" int bmg_model_sample_beta_sigma2(bmg_model* m) {
  if (!m) { set_error("null model"); return BMG_E_ARG; }
  ModelOpArgs a{};
  a.op = 4;
  int rc = launch_model_op(m->b, a);
  if (rc) return rc;
  ChainScal s;
  BMG_HIP(hipMemcpy(&s, m->b->d_scal, sizeof(ChainScal), hipMemcpyDeviceToHost));
  if (s.err & ERR_SIGMA2) { set_error("sigma2 draw did not finish (non-finite residual sum)"); return BMG_E_NUMERIC; }
  return BMG_OK;
}

extern "C" int bmg_model_compute_pve(bmg_model* m, double pves[3]) {
  if (!m || !pves) { set_error("null argument"); return BMG_E_ARG; }
  ModelOpArgs a{};
  a.op = 5;
  int rc = launch_model_op(m->b, a);
  if (rc) return rc;
  ChainScal s;
  BMG_HIP(hipMemcpy(&s, m->b->d_scal, sizeof(ChainScal), hipMemcpyDeviceToHost));
  for (int i = 0; i < 3; ++i) pves[i] = s.pves[i];
  return BMG_OK;
}

extern "C" int bmg_model_beta(bmg_model* m, double* beta, double* sigma2, double* y_hat, int* len) {
  if (!m) { set_error("null model"); return BMG_E_ARG; }
  bmg_batch* b = m->b;
  ChainScal s;
  BMG_HIP(hipMemcpy(&s, b->d_scal, sizeof(ChainScal), hipMemcpyDeviceToHost));
  const ChainArrays& A = b->h_arrays[0];
  if (beta && s.beta_len > 0)
    BMG_HIP(hipMemcpy(beta, A.vec + 3 * BMG_KMAXG, sizeof(double) * s.beta_len, hipMemcpyDeviceToHost));
  if (sigma2) *sigma2 = s.sigma2;
  if (y_hat) BMG_HIP(hipMemcpy(y_hat, A.y_hat, sizeof(double) * b->h_data[0].n, hipMemcpyDeviceToHost));
  if (len) *len = s.beta_len;
  return BMG_OK;
}

extern "C" int bmg_model_rb(bmg_model* m, const double* y_hat, double* p_r) {
  if (!m || !y_hat || !p_r) { set_error("null argument"); return BMG_E_ARG; }
  const bmg_ctx* x = m->b->ctxs[0];
  if (x->n_types != 1) {
    set_error("the RB scan of a model takes one effect type (several: the type softmax is not built)");
    return BMG_E_UNSUPPORTED;
  }
  return bmg_batch_rb(m->b, 0, y_hat, p_r);
}

extern "C" int bmg_model_commit(bmg_model* current, const bmg_model* proposed) {
  return bmg_model_copy(current, proposed);
}

extern "C" int bmg_model_rollback(bmg_model* proposed, const bmg_model* current) {
  return bmg_model_copy(proposed, current);
}

extern "C" int bmg_model_destroy(bmg_model* m) {
  if (!m) return BMG_OK;
  bmg_batch_destroy(m->b);
  delete m;
  return BMG_OK;
}

extern "C" int bmg_model_add_term(bmg_model* m, uint32_t ind, int t_ind, const double* d) {
  if (!m || !d) { set_error("null argument"); return BMG_E_ARG; }
  if (ind >= m->b->h_data[0].p) { set_error("snp out of range"); return BMG_E_ARG; }
  ModelOpArgs a{};
  a.op = 0; a.ind = ind; a.t_ind = t_ind; a.d0 = d[0];
  const DevData& D = m->b->h_data[0];
  a.d1 = (t_ind >= 0 && t_ind < D.n_types && D.types[t_ind] == BMG_AH) ? d[1] : 0.0;
  int rc = launch_model_op(m->b, a);
  if (rc) return rc;
  ChainScal s;
  BMG_HIP(hipMemcpy(&s, m->b->d_scal, sizeof(ChainScal), hipMemcpyDeviceToHost));
  if (s.err & ERR_CAPACITY) { set_error("model capacity exceeded"); return BMG_E_CAPACITY; }
  return BMG_OK;
}

extern "C" int bmg_model_remove_term(bmg_model* m, int model_ind) {
  if (!m) { set_error("null model"); return BMG_E_ARG; }
  ChainScal s;
  BMG_HIP(hipMemcpy(&s, m->b->d_scal, sizeof(ChainScal), hipMemcpyDeviceToHost));
  if (model_ind < 0 || model_ind >= s.nloci) { set_error("model_ind out of range"); return BMG_E_ARG; }
  ModelOpArgs a{};
  a.op = 1; a.model_ind = model_ind;
  return launch_model_op(m->b, a);
}

extern "C" int bmg_model_compute_log_likelihood(bmg_model* m) {
  if (!m) { set_error("null model"); return BMG_E_ARG; }
  ModelOpArgs a{};
  a.op = 2;
  return launch_model_op(m->b, a);
}

extern "C" int bmg_model_swapadj(bmg_model* m, int col) {
  if (!m) { set_error("null model"); return BMG_E_ARG; }
  ChainScal s;
  BMG_HIP(hipMemcpy(&s, m->b->d_scal, sizeof(ChainScal), hipMemcpyDeviceToHost));
  if (col < 0 || col + 1 >= s.cols) { set_error("col out of range"); return BMG_E_ARG; }
  ModelOpArgs a{};
  a.op = 3; a.model_ind = col;
  return launch_model_op(m->b, a);
}

extern "C" int bmg_model_log_likelihood(bmg_model* m, double* out) {
  if (!m || !out) { set_error("null argument"); return BMG_E_ARG; }
  ChainScal s;
  BMG_HIP(hipMemcpy(&s, m->b->d_scal, sizeof(ChainScal), hipMemcpyDeviceToHost));
  *out = s.ll;
  return BMG_OK;
}

extern "C" int bmg_model_size(bmg_model* m, int* out) {
  if (!m || !out) { set_error("null argument"); return BMG_E_ARG; }
  ChainScal s;
  BMG_HIP(hipMemcpy(&s, m->b->d_scal, sizeof(ChainScal), hipMemcpyDeviceToHost));
  *out = s.nloci;
  return BMG_OK;
}

extern "C" int bmg_model_set_d(bmg_model* m, int x_ind, double val) {
  if (!m) { set_error("null model"); return BMG_E_ARG; }
  const int K = m->b->kmax;
  if (x_ind < 0 || x_ind >= K) { set_error("x_ind out of range"); return BMG_E_ARG; }
  BMG_HIP(hipMemcpy(m->b->h_arrays[0].vec + 2 * BMG_KMAXG + x_ind, &val, sizeof(double), hipMemcpyHostToDevice));
  return BMG_OK;
}

extern "C" int bmg_model_copy(bmg_model* dst, const bmg_model* src) {
  if (!dst || !src) { set_error("null model"); return BMG_E_ARG; }
  bmg_batch* a = src->b;
  bmg_batch* b = dst->b;
  if (a->kmax != b->kmax || a->h_data[0].p != b->h_data[0].p) { set_error("incompatible models"); return BMG_E_ARG; }
  BMG_HIP(hipMemcpy(b->pool, a->pool, a->per_chain, hipMemcpyDeviceToDevice));
  // the pool copy moved every array; the scalar block carries the rest
  BMG_HIP(hipMemcpy(b->d_scal, a->d_scal, sizeof(ChainScal), hipMemcpyDeviceToDevice));
  return BMG_OK;
}

extern "C" int bmg_model_get(bmg_model* m, double* xx, double* l, double* xy, double* v, double* d,
                             double scal[6]) {
  if (!m) { set_error("null model"); return BMG_E_ARG; }
  bmg_batch* b = m->b;
  const int K = BMG_KMAXG;   // per-column arrays are laid out for BMG_KMAXG columns
  ChainScal s;
  BMG_HIP(hipMemcpy(&s, b->d_scal, sizeof(ChainScal), hipMemcpyDeviceToHost));
  int k = s.cols;
  const ChainArrays& A = b->h_arrays[0];
  std::vector<double> Lp((size_t)K * (K + 1) / 2), xxs((size_t)K * K), vec(5 * K);
  std::vector<int32_t> idx(8 * K);
  BMG_HIP(hipMemcpy(Lp.data(), A.Lcur, sizeof(double) * Lp.size(), hipMemcpyDeviceToHost));
  BMG_HIP(hipMemcpy(xxs.data(), A.xxs, sizeof(double) * xxs.size(), hipMemcpyDeviceToHost));
  BMG_HIP(hipMemcpy(vec.data(), A.vec, sizeof(double) * vec.size(), hipMemcpyDeviceToHost));
  BMG_HIP(hipMemcpy(idx.data(), A.idx, sizeof(int32_t) * idx.size(), hipMemcpyDeviceToHost));
  for (int cc = 0; cc < k; ++cc)
    for (int r = 0; r < k; ++r) {
      if (xx) xx[cc * k + r] = r <= cc ? xxs[(size_t)idx[3 * K + r] * K + idx[3 * K + cc]] : 0.0;
      if (l) l[cc * k + r] = r <= cc ? Lp[cc * (cc + 1) / 2 + r] : 0.0;
    }
  if (xy) for (int i = 0; i < k; ++i) xy[i] = vec[i];
  if (v) for (int i = 0; i < k; ++i) v[i] = vec[K + i];
  if (d) for (int i = 0; i < k; ++i) d[i] = vec[2 * K + i];
  if (scal) { scal[0] = s.ll; scal[1] = s.syx; scal[2] = s.ldq; scal[3] = s.ldl; scal[4] = s.sigma2; scal[5] = k; }
  return BMG_OK;
}
