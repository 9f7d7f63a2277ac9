// bmg_mlp.hip -- the learned-proposal network on gfx950.
//
// One workgroup (4 waves) owns 64 rows.  Every product is a
// v_mfma_f32_16x16x32_bf16 tile: lane l supplies A[l % 16][8 (l / 16) + i]
// and B[8 (l / 16) + i][l % 16] (i < 8) and receives C[4 (l / 16) + r][l % 16]
// (r < 4).  Wave w owns hidden columns [64 w, 64 w + 64).  Activations live in
// LDS as bf16, in both row-major and transposed form where a later product
// reduces over rows; weights come from L2 as bf16 copies whose reduction
// index is contiguous.  Each workgroup writes its fp32 gradient partials,
// and k_mlp_reduce sums them in workgroup order (the same gradient for the
// same batch, run after run).
#include <hip/hip_runtime.h>
#include <math.h>
#include <string.h>

#include <vector>

#include "bmg_internal.h"
#include <algorithm>
#include "bmg_mlp.h"

namespace bmg {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7F800000u) == 0x7F800000u && (u & 0x007FFFFFu)) return 0x7FC0u;
  u = u + 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ __forceinline__ bf16x8 ld8(const uint16_t* p) {
  uint4 u = *(const uint4*)p;
  bf16x8 r;
  __builtin_memcpy(&r, &u, 16);
  return r;
}
__device__ __forceinline__ bf16x8 zero8() {
  uint4 u = make_uint4(0, 0, 0, 0);
  bf16x8 r;
  __builtin_memcpy(&r, &u, 16);
  return r;
}
__device__ __forceinline__ f4 mfma(bf16x8 a, bf16x8 b, f4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}
// sum over the 16 lanes of a row group (lanes that differ in l % 16)
__device__ __forceinline__ float sum16(float v) {
  v += __shfl_xor(v, 1); v += __shfl_xor(v, 2); v += __shfl_xor(v, 4); v += __shfl_xor(v, 8);
  return v;
}
// sum over the 4 row groups (lanes l, l + 16, l + 32, l + 48)
__device__ __forceinline__ float sum4g(float v) {
  v += __shfl_xor(v, 16); v += __shfl_xor(v, 32);
  return v;
}

// LDS layout (bytes): Xs 2048 | A1s 32768 | outp 1024 | gs 256 | XT 2048 |
// A1T 32768 | dZ2s 32768 | dZ2T 32768
#define SM_XS 0
#define SM_A1S 2048
#define SM_OUTP (SM_A1S + 32768)
#define SM_GS (SM_OUTP + 1024)
#define SM_XT (SM_GS + 256)
#define SM_A1T (SM_XT + 2048)
#define SM_DZ2S (SM_A1T + 32768)
#define SM_DZ2T (SM_DZ2S + 32768)
#define SM_FWD SM_XT
#define SM_TRAIN (SM_DZ2T + 32768)

template <bool TRAIN>
__global__ __launch_bounds__(256) void k_mlp(const uint16_t* __restrict__ X, const float* __restrict__ lab,
                                             int64_t rows, float inv_b, MlpArgs m, float* __restrict__ out) {
  extern __shared__ __align__(16) uint8_t smem[];
  uint16_t* Xs = (uint16_t*)(smem + SM_XS);
  uint16_t* A1s = (uint16_t*)(smem + SM_A1S);
  float* outp = (float*)(smem + SM_OUTP);
  float* gs = (float*)(smem + SM_GS);
  uint16_t* XT = (uint16_t*)(smem + SM_XT);
  uint16_t* A1T = (uint16_t*)(smem + SM_A1T);
  uint16_t* dZ2s = (uint16_t*)(smem + SM_DZ2S);
  uint16_t* dZ2T = (uint16_t*)(smem + SM_DZ2T);
  const int t = threadIdx.x, lane = t & 63, w = t >> 6, lr = lane & 15, grp = lane >> 4;
  const int64_t r0 = (int64_t)blockIdx.x * MLP_ROWS;

  // ---- input tile
  if (t < 128) {
    int row = t >> 1, half = t & 1;
    uint4 v = make_uint4(0, 0, 0, 0);
    if (r0 + row < rows) v = *(const uint4*)(X + (r0 + row) * MLP_IN + 8 * half);
    *(uint4*)(Xs + row * MLP_IN + 8 * half) = v;
    if (TRAIN) {
      uint16_t e[8];
      __builtin_memcpy(e, &v, 16);
#pragma unroll
      for (int i = 0; i < 8; ++i) XT[(8 * half + i) * 64 + row] = e[i];
    }
  }
  __syncthreads();

  // ---- layer 1: z1 = x W1 + b1, a1 = bf16(relu(z1))
#pragma unroll
  for (int nt = 0; nt < 4; ++nt) {
    const int col = 64 * w + 16 * nt + lr;
    bf16x8 bfr = grp < 2 ? ld8(m.W1t + col * MLP_IN + 8 * grp) : zero8();
    const float bb = m.P[MLP_OFF_B1 + 64 * w + 16 * nt + lr];
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) {
      bf16x8 afr = grp < 2 ? ld8(Xs + (16 * mt + lr) * MLP_IN + 8 * grp) : zero8();
      f4 acc = mfma(afr, bfr, f4{0.f, 0.f, 0.f, 0.f});
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        int row = 16 * mt + 4 * grp + r;
        float z = acc[r] + bb;
        uint16_t a = f2bf(z > 0.0f ? z : 0.0f);
        A1s[row * MLP_HID + col] = a;
        if (TRAIN) A1T[col * 64 + row] = a;
      }
    }
  }
  __syncthreads();

  // ---- layer 2: z2 = a1 W2 + b2
  f4 acc[4][4];
#pragma unroll
  for (int mt = 0; mt < 4; ++mt)
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) acc[mt][nt] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll 2
  for (int ks = 0; ks < 8; ++ks) {
    bf16x8 af[4];
#pragma unroll
    for (int mt = 0; mt < 4; ++mt) af[mt] = ld8(A1s + (16 * mt + lr) * MLP_HID + 32 * ks + 8 * grp);
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
      bf16x8 bfr = ld8(m.W2t + (64 * w + 16 * nt + lr) * MLP_HID + 32 * ks + 8 * grp);
#pragma unroll
      for (int mt = 0; mt < 4; ++mt) acc[mt][nt] = mfma(af[mt], bfr, acc[mt][nt]);
    }
  }
  // ---- layer 3: out = relu(z2) . w3 + b3
  float part[4][4];
#pragma unroll
  for (int mt = 0; mt < 4; ++mt)
#pragma unroll
    for (int r = 0; r < 4; ++r) part[mt][r] = 0.0f;
#pragma unroll
  for (int nt = 0; nt < 4; ++nt) {
    const int col = 64 * w + 16 * nt + lr;
    const float bb = m.P[MLP_OFF_B2 + col];
    const float w3f = bf2f(m.w3b[col]);
#pragma unroll
    for (int mt = 0; mt < 4; ++mt)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        float z = acc[mt][nt][r] + bb;
        acc[mt][nt][r] = z;
        part[mt][r] += (z > 0.0f ? z : 0.0f) * w3f;
      }
  }
#pragma unroll
  for (int mt = 0; mt < 4; ++mt)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      float v = sum16(part[mt][r]);
      if (lr == 0) outp[w * 64 + 16 * mt + 4 * grp + r] = v;
    }
  __syncthreads();
  if (t < 64) {
    float s = outp[t] + outp[64 + t] + outp[128 + t] + outp[192 + t] + m.P[MLP_OFF_B3];
    bool valid = r0 + t < rows;
    if (out && valid) out[r0 + t] = s;
    if (TRAIN) {
      float g = 0.0f;
      if (valid) {
        float e = s - lab[r0 + t];
        g = (e > 1.0f ? 1.0f : (e < -1.0f ? -1.0f : e)) * inv_b;
      }
      gs[t] = g;
    }
  }
  if (!TRAIN) return;
  __syncthreads();
  // this workgroup's gradient: every parameter is written exactly once below
  float* gp = m.GP + (size_t)(blockIdx.x % MLP_GP_WG) * MLP_NPAR;

  // ---- backward through layer 3: dw3, db2, dz2 = g w3 [z2 > 0]
  if (t == 0) {
    float s = 0.0f;
    for (int i = 0; i < 64; ++i) s += gs[i];
    gp[MLP_OFF_B3] = s;
  }
#pragma unroll
  for (int nt = 0; nt < 4; ++nt) {
    const int col = 64 * w + 16 * nt + lr;
    const float w3f = bf2f(m.w3b[col]);
    float pw3 = 0.0f, pb2 = 0.0f;
#pragma unroll
    for (int mt = 0; mt < 4; ++mt)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        int row = 16 * mt + 4 * grp + r;
        float z = acc[mt][nt][r];
        float g = gs[row];
        pw3 += g * (z > 0.0f ? z : 0.0f);
        float dz = z > 0.0f ? g * w3f : 0.0f;
        pb2 += dz;
        uint16_t d = f2bf(dz);
        dZ2s[row * MLP_HID + col] = d;
        dZ2T[col * 64 + row] = d;
      }
    pw3 = sum4g(pw3);
    pb2 = sum4g(pb2);
    if (grp == 0) {
      gp[MLP_OFF_W3 + col] = pw3;
      gp[MLP_OFF_B2 + col] = pb2;
    }
  }
  __syncthreads();

  // ---- dW2 = a1^T dz2 (rows h of wave w, columns g in two halves)
#pragma unroll
  for (int half = 0; half < 2; ++half) {
    f4 a2[4][8];
#pragma unroll
    for (int mt = 0; mt < 4; ++mt)
#pragma unroll
      for (int nt = 0; nt < 8; ++nt) a2[mt][nt] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int ks = 0; ks < 2; ++ks) {
      bf16x8 af[4];
#pragma unroll
      for (int mt = 0; mt < 4; ++mt) af[mt] = ld8(A1T + (64 * w + 16 * mt + lr) * 64 + 32 * ks + 8 * grp);
#pragma unroll
      for (int nt = 0; nt < 8; ++nt) {
        bf16x8 bfr = ld8(dZ2T + (128 * half + 16 * nt + lr) * 64 + 32 * ks + 8 * grp);
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) a2[mt][nt] = mfma(af[mt], bfr, a2[mt][nt]);
      }
    }
#pragma unroll
    for (int mt = 0; mt < 4; ++mt)
#pragma unroll
      for (int nt = 0; nt < 8; ++nt)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          int h = 64 * w + 16 * mt + 4 * grp + r;
          int g = 128 * half + 16 * nt + lr;
          gp[MLP_OFF_W2 + h * MLP_HID + g] = a2[mt][nt][r];
        }
  }
  __syncthreads();   // A1T is reused for dz1 below

  // ---- dA1 = dz2 W2^T, dz1 = dA1 [a1 > 0], db1
  {
    f4 a3[4][4];
#pragma unroll
    for (int mt = 0; mt < 4; ++mt)
#pragma unroll
      for (int nt = 0; nt < 4; ++nt) a3[mt][nt] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll 2
    for (int ks = 0; ks < 8; ++ks) {
      bf16x8 af[4];
#pragma unroll
      for (int mt = 0; mt < 4; ++mt) af[mt] = ld8(dZ2s + (16 * mt + lr) * MLP_HID + 32 * ks + 8 * grp);
#pragma unroll
      for (int nt = 0; nt < 4; ++nt) {
        bf16x8 bfr = ld8(m.W2b + (64 * w + 16 * nt + lr) * MLP_HID + 32 * ks + 8 * grp);
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) a3[mt][nt] = mfma(af[mt], bfr, a3[mt][nt]);
      }
    }
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
      const int h = 64 * w + 16 * nt + lr;
      float pb1 = 0.0f;
#pragma unroll
      for (int mt = 0; mt < 4; ++mt)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          int row = 16 * mt + 4 * grp + r;
          float dz1 = A1s[row * MLP_HID + h] != 0 ? a3[mt][nt][r] : 0.0f;
          pb1 += dz1;
          A1T[h * 64 + row] = f2bf(dz1);
        }
      pb1 = sum4g(pb1);
      if (grp == 0) gp[MLP_OFF_B1 + h] = pb1;
    }
  }
  __syncthreads();

  // ---- dW1 = x^T dz1 (16 input rows x 256)
#pragma unroll
  for (int nt = 0; nt < 4; ++nt) {
    const int h = 64 * w + 16 * nt + lr;
    f4 a4 = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int ks = 0; ks < 2; ++ks) {
      bf16x8 af = ld8(XT + lr * 64 + 32 * ks + 8 * grp);
      bf16x8 bfr = ld8(A1T + (64 * w + 16 * nt + lr) * 64 + 32 * ks + 8 * grp);
      a4 = mfma(af, bfr, a4);
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      int k = 4 * grp + r;
      gp[MLP_OFF_W1 + k * MLP_HID + (64 * w + 16 * nt) + lr] = a4[r];
    }
    (void)h;
  }
}

// bf16 copies of one parameter (index i of the master array)
__device__ __forceinline__ void write_copy(const MlpArgs& m, int i, float p) {
  uint16_t b = f2bf(p);
  if (i < MLP_OFF_B1) {
    int k = i / MLP_HID, h = i % MLP_HID;
    m.W1t[h * MLP_IN + k] = b;
  } else if (i >= MLP_OFF_W2 && i < MLP_OFF_B2) {
    int e = i - MLP_OFF_W2, h = e / MLP_HID, g = e % MLP_HID;
    m.W2t[g * MLP_HID + h] = b;
    m.W2b[h * MLP_HID + g] = b;
  } else if (i >= MLP_OFF_W3 && i < MLP_OFF_B3) {
    m.w3b[i - MLP_OFF_W3] = b;
  }
}

// G[i] += sum over the launch's workgroups of their partials, in order
__global__ void k_mlp_reduce(MlpArgs m, int nwg) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= MLP_NPAR) return;
  float acc = 0.0f;
  for (int b = 0; b < nwg; ++b) acc += m.GP[(size_t)b * MLP_NPAR + i];
  m.G[i] += acc;
}

__global__ void k_mlp_copies(MlpArgs m) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < MLP_NPAR) write_copy(m, i, m.P[i]);
}

// Adam (Kingma & Ba 2015), fp32, beta1 0.9, beta2 0.999, eps 1e-8
__global__ void k_mlp_adam(MlpArgs m, float lr, float c1, float c2) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= MLP_NPAR) return;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  float g = m.G[i];
  m.G[i] = 0.0f;
  float mm = b1 * m.M[i] + (1.0f - b1) * g;
  float vv = b2 * m.V[i] + (1.0f - b2) * g * g;
  m.M[i] = mm; m.V[i] = vv;
  float mh = mm / c1, vh = vv / c2;
  float p = m.P[i] - lr * mh / (sqrtf(vh) + eps);
  m.P[i] = p;
  write_copy(m, i, p);
}

__global__ void k_fill_f32(float* p, int64_t n, float v) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

uint16_t bf16_bits(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  if ((u & 0x7F800000u) == 0x7F800000u && (u & 0x007FFFFFu)) return 0x7FC0u;
  u = u + 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

int mlp_alloc(MlpArgs* m) {
  *m = MlpArgs{};
  BMG_HIP(hipMalloc(&m->P, sizeof(float) * MLP_NPAR));
  BMG_HIP(hipMalloc(&m->G, sizeof(float) * MLP_NPAR));
  BMG_HIP(hipMalloc(&m->M, sizeof(float) * MLP_NPAR));
  BMG_HIP(hipMalloc(&m->V, sizeof(float) * MLP_NPAR));
  BMG_HIP(hipMalloc(&m->GP, sizeof(float) * MLP_NPAR * MLP_GP_WG));
  BMG_HIP(hipMalloc(&m->W1t, sizeof(uint16_t) * MLP_HID * MLP_IN));
  BMG_HIP(hipMalloc(&m->W2t, sizeof(uint16_t) * MLP_HID * MLP_HID));
  BMG_HIP(hipMalloc(&m->W2b, sizeof(uint16_t) * MLP_HID * MLP_HID));
  BMG_HIP(hipMalloc(&m->w3b, sizeof(uint16_t) * MLP_HID));
  unsigned grid = (MLP_NPAR + 255) / 256;
  float* bufs[4] = {m->P, m->G, m->M, m->V};
  for (float* b : bufs) hipLaunchKernelGGL(k_fill_f32, dim3(grid), dim3(256), 0, 0, b, (int64_t)MLP_NPAR, 0.0f);
  BMG_HIP(hipGetLastError());
  BMG_HIP(hipDeviceSynchronize());
  return mlp_launch_copies(*m, 0);
}

void mlp_free(MlpArgs* m) {
  hipFree(m->P); hipFree(m->G); hipFree(m->M); hipFree(m->V); hipFree(m->GP);
  hipFree(m->W1t); hipFree(m->W2t); hipFree(m->W2b); hipFree(m->w3b);
  *m = MlpArgs{};
}

int mlp_launch_copies(const MlpArgs& m, hipStream_t s) {
  hipLaunchKernelGGL(k_mlp_copies, dim3((MLP_NPAR + 255) / 256), dim3(256), 0, s, m);
  BMG_HIP(hipGetLastError());
  return BMG_OK;
}

int mlp_launch_forward(const MlpArgs& m, const uint16_t* X, int64_t rows, float* out, hipStream_t s) {
  if (rows <= 0) return BMG_OK;
  unsigned grid = (unsigned)((rows + MLP_ROWS - 1) / MLP_ROWS);
  hipLaunchKernelGGL(k_mlp<false>, dim3(grid), dim3(256), SM_FWD, s, X, (const float*)nullptr, rows, 0.0f, m,
                     out);
  BMG_HIP(hipGetLastError());
  return BMG_OK;
}

int mlp_launch_train(const MlpArgs& m, const uint16_t* X, const float* labels, int64_t rows, float* out,
                     hipStream_t s) {
  if (rows <= 0) return BMG_OK;
  // launches of at most MLP_GP_WG workgroups, each reduced in order into G;
  // every launch scales by the whole batch, so G is the mean's gradient
  const float inv_b = 1.0f / (float)rows;
  const int64_t chunk = (int64_t)MLP_GP_WG * MLP_ROWS;
  for (int64_t r0 = 0; r0 < rows; r0 += chunk) {
    const int64_t nr = std::min<int64_t>(chunk, rows - r0);
    const unsigned grid = (unsigned)((nr + MLP_ROWS - 1) / MLP_ROWS);
    hipLaunchKernelGGL(k_mlp<true>, dim3(grid), dim3(256), SM_TRAIN, s, X + r0 * MLP_IN, labels + r0, nr, inv_b, m,
                       out ? out + r0 : nullptr);
    hipLaunchKernelGGL(k_mlp_reduce, dim3((MLP_NPAR + 255) / 256), dim3(256), 0, s, m, (int)grid);
  }
  BMG_HIP(hipGetLastError());
  return BMG_OK;
}

int mlp_launch_adam(const MlpArgs& m, float lr, int t, hipStream_t s) {
  const float c1 = 1.0f - powf(0.9f, (float)t), c2 = 1.0f - powf(0.999f, (float)t);
  hipLaunchKernelGGL(k_mlp_adam, dim3((MLP_NPAR + 255) / 256), dim3(256), 0, s, m, lr, c1, c2);
  BMG_HIP(hipGetLastError());
  return BMG_OK;
}

}  // namespace bmg

// ------------------------------------------------------------ C ABI -----
using namespace bmg;

struct bmg_mlp {
  int device = 0;
  MlpArgs m{};
  hipStream_t s = nullptr;
  int t = 0;
};

static int upload_x(const float* X, int64_t rows, uint16_t** Xd) {
  std::vector<uint16_t> h((size_t)rows * MLP_IN);
  for (size_t i = 0; i < h.size(); ++i) h[i] = bf16_bits(X[i]);
  BMG_HIP(hipMalloc(Xd, sizeof(uint16_t) * h.size()));
  BMG_HIP(hipMemcpy(*Xd, h.data(), sizeof(uint16_t) * h.size(), hipMemcpyHostToDevice));
  return BMG_OK;
}

extern "C" int bmg_mlp_nparams(void) { return MLP_NPAR; }

extern "C" int bmg_mlp_create(int device, bmg_mlp** out) {
  if (!out) { set_error("null argument"); return BMG_E_ARG; }
  *out = nullptr;
  BMG_HIP(hipSetDevice(device));
  bmg_mlp* n = new bmg_mlp();
  n->device = device;
  int rc = mlp_alloc(&n->m);
  if (rc) { mlp_free(&n->m); delete n; return rc; }
  hipError_t e = hipStreamCreateWithFlags(&n->s, hipStreamNonBlocking);
  if (e != hipSuccess) { mlp_free(&n->m); delete n; return hip_fail(e, "hipStreamCreate"); }
  *out = n;
  return BMG_OK;
}

extern "C" int bmg_mlp_destroy(bmg_mlp* n) {
  if (!n) return BMG_OK;
  hipSetDevice(n->device);
  if (n->s) { hipStreamSynchronize(n->s); hipStreamDestroy(n->s); }
  mlp_free(&n->m);
  delete n;
  return BMG_OK;
}

extern "C" int bmg_mlp_set_params(bmg_mlp* n, const float* params) {
  if (!n || !params) { set_error("null argument"); return BMG_E_ARG; }
  BMG_HIP(hipSetDevice(n->device));
  BMG_HIP(hipMemcpy(n->m.P, params, sizeof(float) * MLP_NPAR, hipMemcpyHostToDevice));
  unsigned grid = (MLP_NPAR + 255) / 256;
  float* bufs[3] = {n->m.G, n->m.M, n->m.V};
  for (float* b : bufs) hipLaunchKernelGGL(k_fill_f32, dim3(grid), dim3(256), 0, n->s, b, (int64_t)MLP_NPAR, 0.0f);
  BMG_HIP(hipGetLastError());
  int rc = mlp_launch_copies(n->m, n->s);
  if (rc) return rc;
  n->t = 0;
  BMG_HIP(hipStreamSynchronize(n->s));
  return BMG_OK;
}

extern "C" int bmg_mlp_get_params(bmg_mlp* n, float* params) {
  if (!n || !params) { set_error("null argument"); return BMG_E_ARG; }
  BMG_HIP(hipSetDevice(n->device));
  BMG_HIP(hipStreamSynchronize(n->s));
  BMG_HIP(hipMemcpy(params, n->m.P, sizeof(float) * MLP_NPAR, hipMemcpyDeviceToHost));
  return BMG_OK;
}

extern "C" int bmg_mlp_forward(bmg_mlp* n, const float* X, int64_t rows, float* out) {
  if (!n || !X || !out || rows < 1) { set_error("bad argument"); return BMG_E_ARG; }
  BMG_HIP(hipSetDevice(n->device));
  uint16_t* Xd = nullptr;
  float* od = nullptr;
  int rc = upload_x(X, rows, &Xd);
  if (!rc && hipMalloc(&od, sizeof(float) * rows) != hipSuccess) { set_error("hipMalloc"); rc = BMG_E_HIP; }
  if (!rc) rc = mlp_launch_forward(n->m, Xd, rows, od, n->s);
  if (!rc && hipStreamSynchronize(n->s) != hipSuccess) { set_error("forward failed"); rc = BMG_E_HIP; }
  if (!rc && hipMemcpy(out, od, sizeof(float) * rows, hipMemcpyDeviceToHost) != hipSuccess) {
    set_error("copy failed"); rc = BMG_E_HIP;
  }
  hipFree(Xd); hipFree(od);
  return rc;
}

static int train_common(bmg_mlp* n, const float* X, const float* labels, int64_t rows) {
  uint16_t* Xd = nullptr;
  float* ld = nullptr;
  int rc = upload_x(X, rows, &Xd);
  if (!rc && hipMalloc(&ld, sizeof(float) * rows) != hipSuccess) { set_error("hipMalloc"); rc = BMG_E_HIP; }
  if (!rc && hipMemcpy(ld, labels, sizeof(float) * rows, hipMemcpyHostToDevice) != hipSuccess) {
    set_error("copy failed"); rc = BMG_E_HIP;
  }
  if (!rc) rc = mlp_launch_train(n->m, Xd, ld, rows, nullptr, n->s);
  if (!rc && hipStreamSynchronize(n->s) != hipSuccess) { set_error("train kernel failed"); rc = BMG_E_HIP; }
  hipFree(Xd); hipFree(ld);
  return rc;
}

extern "C" int bmg_mlp_gradients(bmg_mlp* n, const float* X, const float* labels, int64_t rows, float* grads) {
  if (!n || !X || !labels || !grads || rows < 1) { set_error("bad argument"); return BMG_E_ARG; }
  BMG_HIP(hipSetDevice(n->device));
  int rc = train_common(n, X, labels, rows);
  if (rc) return rc;
  BMG_HIP(hipMemcpy(grads, n->m.G, sizeof(float) * MLP_NPAR, hipMemcpyDeviceToHost));
  hipLaunchKernelGGL(k_fill_f32, dim3((MLP_NPAR + 255) / 256), dim3(256), 0, n->s, n->m.G, (int64_t)MLP_NPAR, 0.0f);
  BMG_HIP(hipGetLastError());
  BMG_HIP(hipStreamSynchronize(n->s));
  return BMG_OK;
}

extern "C" int bmg_mlp_adam_step(bmg_mlp* n, const float* X, const float* labels, int64_t rows, float lr) {
  if (!n || !X || !labels || rows < 1) { set_error("bad argument"); return BMG_E_ARG; }
  BMG_HIP(hipSetDevice(n->device));
  int rc = train_common(n, X, labels, rows);
  if (rc) return rc;
  n->t += 1;
  rc = mlp_launch_adam(n->m, lr, n->t, n->s);
  if (rc) return rc;
  BMG_HIP(hipStreamSynchronize(n->s));
  return BMG_OK;
}
