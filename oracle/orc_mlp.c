/*
 * orc_mlp.c -- CPU oracle of the learned-proposal network.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  The network has no counterpart
 * in the reference; this file states its arithmetic once, with the same
 * rounding points as the device kernels, so that the device forward pass
 * and gradients can be checked:
 *
 *   x   = bf16(features)                                (16)
 *   z1  = x W1 + b1,  a1 = bf16(relu(z1))               (256)
 *   z2  = a1 bf16(W2) + b2,  a2 = relu(z2)              (256, fp32)
 *   out = a2 . bf16(w3) + b3
 *   loss = mean over rows of huber_1(out - label)
 *   g   = clamp(out - label, -1, 1) / B
 *   dz2 = g bf16(w3) [z2 > 0],  dz2b = bf16(dz2)
 *   dz1 = (dz2b bf16(W2)^T) [a1 > 0],  dz1b = bf16(dz1)
 *   dW2 = a1^T dz2b, dW1 = x^T dz1b, db2 = sum dz2, db1 = sum dz1,
 *   dw3 = sum g a2, db3 = sum g
 *
 * Sums are taken in double here; the device accumulates in fp32 on the
 * MFMA units, so the two agree to fp32 rounding, not bit for bit.
 */
#include "oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define IN ORC_MLP_IN
#define HID ORC_MLP_HID

/* round-to-nearest-even float -> bfloat16 -> float */
float orc_bf16_round(float f)
{
  uint32_t u;
  memcpy(&u, &f, 4);
  if ((u & 0x7F800000u) == 0x7F800000u && (u & 0x007FFFFFu)) {
    u = 0x7FC00000u;
  } else {
    u = u + 0x7FFFu + ((u >> 16) & 1u);
    u &= 0xFFFF0000u;
  }
  float r;
  memcpy(&r, &u, 4);
  return r;
}

typedef struct {
  const float *W1, *b1, *W2, *b2, *w3, *b3;
} prm;

static prm split(const float* p)
{
  prm q;
  q.W1 = p + ORC_MLP_OFF_W1; q.b1 = p + ORC_MLP_OFF_B1;
  q.W2 = p + ORC_MLP_OFF_W2; q.b2 = p + ORC_MLP_OFF_B2;
  q.w3 = p + ORC_MLP_OFF_W3; q.b3 = p + ORC_MLP_OFF_B3;
  return q;
}

/* one row: fills z1, a1 (bf16 values), z2, a2; returns out */
static double row_forward(const prm* q, const float* xr, float* a1, float* z2,
                          float* a2, float* W1r, float* W2r, float* w3r)
{
  float x[IN];
  for (int k = 0; k < IN; ++k) x[k] = orc_bf16_round(xr[k]);
  for (int h = 0; h < HID; ++h) {
    double s = 0;
    for (int k = 0; k < IN; ++k) s += (double)x[k] * (double)W1r[k * HID + h];
    float z = (float)(s + (double)q->b1[h]);
    a1[h] = orc_bf16_round(z > 0.0f ? z : 0.0f);
  }
  double out = 0;
  for (int g = 0; g < HID; ++g) {
    double s = 0;
    for (int h = 0; h < HID; ++h) s += (double)a1[h] * (double)W2r[h * HID + g];
    float z = (float)(s + (double)q->b2[g]);
    z2[g] = z;
    a2[g] = z > 0.0f ? z : 0.0f;
    out += (double)a2[g] * (double)w3r[g];
  }
  return out + (double)q->b3[0];
}

/* bf16 copies of the weights the device multiplies with */
static void rounded(const prm* q, float* W1r, float* W2r, float* w3r)
{
  for (int i = 0; i < IN * HID; ++i) W1r[i] = orc_bf16_round(q->W1[i]);
  for (int i = 0; i < HID * HID; ++i) W2r[i] = orc_bf16_round(q->W2[i]);
  for (int i = 0; i < HID; ++i) w3r[i] = orc_bf16_round(q->w3[i]);
}

void orc_mlp_forward(const float* params, const float* X, int64_t B, float* out)
{
  prm q = split(params);
  float* W1r = (float*)malloc(sizeof(float) * IN * HID);
  float* W2r = (float*)malloc(sizeof(float) * HID * HID);
  float w3r[HID], a1[HID], z2[HID], a2[HID];
  rounded(&q, W1r, W2r, w3r);
  for (int64_t r = 0; r < B; ++r)
    out[r] = (float)row_forward(&q, X + r * IN, a1, z2, a2, W1r, W2r, w3r);
  free(W1r); free(W2r);
}

double orc_mlp_gradients(const float* params, const float* X, const float* labels, int64_t B,
                         double* grads)
{
  prm q = split(params);
  float* W1r = (float*)malloc(sizeof(float) * IN * HID);
  float* W2r = (float*)malloc(sizeof(float) * HID * HID);
  float w3r[HID], a1[HID], z2[HID], a2[HID], dz2b[HID];
  rounded(&q, W1r, W2r, w3r);
  memset(grads, 0, sizeof(double) * ORC_MLP_NPAR);
  double* gW1 = grads + ORC_MLP_OFF_W1; double* gb1 = grads + ORC_MLP_OFF_B1;
  double* gW2 = grads + ORC_MLP_OFF_W2; double* gb2 = grads + ORC_MLP_OFF_B2;
  double* gw3 = grads + ORC_MLP_OFF_W3; double* gb3 = grads + ORC_MLP_OFF_B3;
  double loss = 0;
  for (int64_t r = 0; r < B; ++r) {
    double out = row_forward(&q, X + r * IN, a1, z2, a2, W1r, W2r, w3r);
    float outf = (float)out;
    float e = outf - labels[r];
    double ae = fabs((double)e);
    loss += ae <= 1.0 ? 0.5 * (double)e * (double)e : ae - 0.5;
    float g = (e > 1.0f ? 1.0f : (e < -1.0f ? -1.0f : e)) / (float)B;
    gb3[0] += g;
    for (int c = 0; c < HID; ++c) {
      gw3[c] += (double)g * (double)a2[c];
      float dz = z2[c] > 0.0f ? g * w3r[c] : 0.0f;
      gb2[c] += dz;
      dz2b[c] = orc_bf16_round(dz);
    }
    for (int h = 0; h < HID; ++h) {
      if (a1[h] != 0.0f)
        for (int c = 0; c < HID; ++c) gW2[h * HID + c] += (double)a1[h] * (double)dz2b[c];
    }
    float x[IN];
    for (int k = 0; k < IN; ++k) x[k] = orc_bf16_round(X[r * IN + k]);
    for (int h = 0; h < HID; ++h) {
      double s = 0;
      for (int c = 0; c < HID; ++c) s += (double)dz2b[c] * (double)W2r[h * HID + c];
      float dz1 = a1[h] > 0.0f ? (float)s : 0.0f;
      gb1[h] += dz1;
      float dz1b = orc_bf16_round(dz1);
      for (int k = 0; k < IN; ++k) gW1[k * HID + h] += (double)x[k] * (double)dz1b;
    }
  }
  free(W1r); free(W2r);
  return loss / (double)B;
}

/* exact-arithmetic variant (no bf16 rounding anywhere): the function whose
 * derivative the gradients approximate, for finite-difference checks */
double orc_mlp_loss_exact(const float* params, const float* X, const float* labels, int64_t B,
                          double* grads)
{
  prm q = split(params);
  double loss = 0;
  if (grads) memset(grads, 0, sizeof(double) * ORC_MLP_NPAR);
  double z1[HID], a1[HID], z2[HID], a2[HID], dz2[HID];
  for (int64_t r = 0; r < B; ++r) {
    const float* x = X + r * IN;
    for (int h = 0; h < HID; ++h) {
      double s = q.b1[h];
      for (int k = 0; k < IN; ++k) s += (double)x[k] * q.W1[k * HID + h];
      z1[h] = s; a1[h] = s > 0 ? s : 0;
    }
    double out = q.b3[0];
    for (int c = 0; c < HID; ++c) {
      double s = q.b2[c];
      for (int h = 0; h < HID; ++h) s += a1[h] * q.W2[h * HID + c];
      z2[c] = s; a2[c] = s > 0 ? s : 0;
      out += a2[c] * q.w3[c];
    }
    double e = out - labels[r], ae = fabs(e);
    loss += ae <= 1.0 ? 0.5 * e * e : ae - 0.5;
    if (!grads) continue;
    double g = (e > 1 ? 1 : (e < -1 ? -1 : e)) / (double)B;
    grads[ORC_MLP_OFF_B3] += g;
    for (int c = 0; c < HID; ++c) {
      grads[ORC_MLP_OFF_W3 + c] += g * a2[c];
      dz2[c] = z2[c] > 0 ? g * q.w3[c] : 0;
      grads[ORC_MLP_OFF_B2 + c] += dz2[c];
    }
    for (int h = 0; h < HID; ++h) {
      double s = 0;
      for (int c = 0; c < HID; ++c) {
        grads[ORC_MLP_OFF_W2 + h * HID + c] += a1[h] * dz2[c];
        s += dz2[c] * q.W2[h * HID + c];
      }
      double dz1 = z1[h] > 0 ? s : 0;
      grads[ORC_MLP_OFF_B1 + h] += dz1;
      for (int k = 0; k < IN; ++k) grads[ORC_MLP_OFF_W1 + k * HID + h] += (double)x[k] * dz1;
    }
  }
  return loss / (double)B;
}

/* Adam (Kingma & Ba 2015) on fp32 master weights; t counts from 1 */
void orc_mlp_adam(float* params, float* m, float* v, const float* grads, float lr, int t)
{
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const float c1 = 1.0f - powf(b1, (float)t), c2 = 1.0f - powf(b2, (float)t);
  for (int i = 0; i < ORC_MLP_NPAR; ++i) {
    float g = grads[i];
    m[i] = b1 * m[i] + (1.0f - b1) * g;
    v[i] = b2 * v[i] + (1.0f - b2) * g * g;
    float mh = m[i] / c1, vh = v[i] / c2;
    params[i] -= lr * mh / (sqrtf(vh) + eps);
  }
}
