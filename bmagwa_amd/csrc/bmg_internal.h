// bmg_internal.h -- host-side structs shared by the engine's translation
// units (not part of the C ABI).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>
#include <vector>

#include "../../include/bmagwa_hip.h"

struct bmg_ctx {
  int device = 0;
  uint32_t n = 0;
  uint64_t p = 0;
  uint32_t W32 = 0;      // 32-bit words per SNP (multiple of 32: 128-byte rows)
  uint32_t n_pad = 0;    // W32 * 16 individuals
  uint32_t* g = nullptr; // p * W32 packed words, SNP-major
  double* y = nullptr;   // n_pad, zero padded
  double* pre = nullptr; // p * offset self moments
  int offset = 0;
  int offset_type[4] = {-1, -1, -1, -1};
  int n_types = 0;
  int types[5] = {0, 0, 0, 0, 0};
  int allow_types[5] = {0, 0, 0, 0, 0};
  int allow_terms[5] = {0, 0, 0, 0, 0};
  double yy = 0, var_y = 0, var_x = 0, mean_x = 0;
  uint64_t n_missing = 0;
  std::vector<double> y_host;
  // covariates (the reference's E without the constant): n x m_e host copy
  // (column-major) and m_e x n_pad on the device, zero past n
  uint32_t m_e = 0;
  std::vector<double> e_host;
  double* E = nullptr;
  hipStream_t stream = nullptr;
  // scan scratch (grown on demand)
  int8_t* dig = nullptr; size_t dig_bytes = 0;
  int32_t* acc = nullptr; size_t acc_bytes = 0;
  double* scale = nullptr;
  double* tmp_r = nullptr; size_t tmp_r_bytes = 0;
  double* tmp_out = nullptr; size_t tmp_out_bytes = 0;
  // scan shapes picked by the autotuner, per number of digit columns
  int scan_nt[17] = {0}, scan_wv[17] = {0};
  int no_tune = 0;  // inside a graph capture: no timed tuning (default shape)
  uint32_t scan_tw[17] = {0};
  unsigned scan_gx[17] = {0};
};

namespace bmg {
void set_error(const std::string& msg);
int hip_fail(hipError_t e, const char* what);

// kernels launched from the C ABI (bmg_data.hip)
int ctx_prepare(bmg_ctx* c, int recode);
int ctx_self_moments(bmg_ctx* c);
int launch_decode(bmg_ctx* c, const uint64_t* snps_dev, uint64_t count, int coding,
                  double* out_dev, hipStream_t s);
// genotypes of a subset of a pool dataset's individuals into c (same p)
int launch_subset(const bmg_ctx* pool, const uint32_t* sub_dev, bmg_ctx* c, hipStream_t s);
int launch_synth(bmg_ctx* c, uint32_t seed, const double* thresh_dev, const double* rho_dev,
                 const uint64_t* blk_start_dev, uint64_t n_blocks, hipStream_t s);
int launch_scan(bmg_ctx* c, const double* r_dev, int nvec, int coding, int digits,
                double* out_dev, hipStream_t s);
// the same scan with its scratch in a file-scope device array (no allocation)
int launch_scan_noalloc(bmg_ctx* c, const double* r_dev, int nvec, int coding, int digits,
                        double* out_dev, hipStream_t s);
int launch_scan_timed(bmg_ctx* c, const double* r_dev, int nvec, int coding, int digits,
                      double* out_dev, int reps, double* ms, hipStream_t s);
int launch_band(bmg_ctx* c, uint32_t W, int ca, int cb, int32_t* out_dev, hipStream_t s);
int launch_band_timed(bmg_ctx* c, uint32_t W, int ca, int cb, int32_t* out_dev, int reps, double* ms,
                      hipStream_t s);
int launch_copy_timed(const void* src, void* dst, uint64_t bytes, int reps, double* ms, hipStream_t s);
int launch_read_timed(const void* src, uint64_t bytes, void* sink, int reps, double* ms, hipStream_t s);
int ensure(void** ptr, size_t* have, size_t need);
}  // namespace bmg

#define BMG_HIP(call)                                     \
  do {                                                    \
    hipError_t _e = (call);                               \
    if (_e != hipSuccess) return bmg::hip_fail(_e, #call); \
  } while (0)
