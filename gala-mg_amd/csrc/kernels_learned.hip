// Learned frontier: a 6 -> 64 -> 1 MLP that predicts, per vertex, whether the
// next decide-and-move sweep will move it.  BF16 operands on the gfx950 MFMA
// units (v_mfma_f32_16x16x16_bf16) with FP32 accumulation; FP32 master weights
// updated by Adam.  No counterpart in the reference (the nearest analogue is
// its sound prune, kernel_functions.cu:944,1028): the plain full sweep has the
// last word and certifies every level before it is accepted.
//
// MFMA 16x16x16 operand layout used throughout (64 lanes, l = lane):
//   A (16 x 16, rows x k):  lane l holds A[l % 16][4 (l / 16) + i], i = 0..3
//   B (16 x 16, k x cols):  lane l holds B[4 (l / 16) + i][l % 16]
//   D (16 x 16):            lane l holds D[4 (l / 16) + i][l % 16]
// Forward: rows = 16 vertices, k = features (6 used, 10 zero), cols = 16 of
// the 64 hidden units (4 MFMAs).  The hidden-gradient tile dH[sample][hidden]
// comes out in the D layout, which is exactly the B layout of the weight
// gradient dW1 = X^T dH (rows = features, k = samples), so backward needs no
// data movement between the two MFMAs.
#include "gala_kernels.h"
#include "gala_device.h"

namespace gala {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ short to_bf16(float x) {  // round to nearest even
  unsigned u = __float_as_uint(x);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (short)(u >> 16);
}

__device__ __forceinline__ float sigmoidf(float z) { return 1.0f / (1.0f + __expf(-z)); }

// Feature vector of vertex v before sweep `it` (see the header comment of
// gala_kernels.h for the definitions).
template <class T>
__device__ __forceinline__ void vertex_features(const DevGraph<T>& g, const DevState<T>& st, int v, int it,
                                                float f[MLP_IN]) {
  const int64_t d = g.end[v] - g.off[v];
  const double K = (double)st.K[v];
  f[0] = __log2f(1.0f + (float)d) * (1.0f / 16.0f);
  f[1] = K > 0 ? (float)((double)st.In[v] / K) : 0.0f;
  const float lg = st.last_gain[v];
  f[2] = lg < -1.0f ? -1.0f : (lg > 1.0f ? 1.0f : lg);
  f[3] = st.moved_it[v] == it - 1 ? 1.0f : 0.0f;
  // the size of its community on the scale of the graph: log(1 + size) /
  // log(1 + n), the same range for graphs of every size
  const int sz = st.size[st.cur[v]];
  f[4] = g.n > 1 ? __log2f(1.0f + (float)sz) / __log2f(1.0f + (float)g.n) : 0.0f;
  f[5] = (float)(it < 16 ? it : 16) * (1.0f / 16.0f);
}

// 16 rows x 64 hidden forward through the MFMA units for the wave's rows.
// x[j] = feature 4*(l/16)+j of row l%16 (zero beyond MLP_IN).  Returns the
// pre-activations acc[t][i] = z1[row 4(l/16)+i][hidden 16 t + l%16].
__device__ __forceinline__ void mlp_hidden(const float x[4], const short* __restrict__ w1b, f32x4 acc[4]) {
  const int lane = threadIdx.x & 63;
  bf16x4 a;
  for (int j = 0; j < 4; j++) a[j] = to_bf16(x[j]);
#pragma unroll
  for (int t = 0; t < 4; t++) {
    bf16x4 b;
#pragma unroll
    for (int j = 0; j < 4; j++) b[j] = w1b[(4 * (lane >> 4) + j) * MLP_H + 16 * t + (lane & 15)];
    f32x4 z = {0.f, 0.f, 0.f, 0.f};
    acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, z, 0, 0, 0);
  }
}

// logits of rows 4(l/16)+i (identical in all 16 lanes of a row group)
__device__ __forceinline__ void mlp_logits(const f32x4 acc[4], const float* __restrict__ prm, float z[4]) {
  const int lane = threadIdx.x & 63;
  const float* b1 = prm + MLP_B1;
  const float* w2 = prm + MLP_W2;
#pragma unroll
  for (int i = 0; i < 4; i++) {
    float s = 0.f;
#pragma unroll
    for (int t = 0; t < 4; t++) {
      const int h = 16 * t + (lane & 15);
      const float a = acc[t][i] + b1[h];
      s += (a > 0.f ? a : 0.f) * w2[h];
    }
#pragma unroll
    for (int k = 1; k < 16; k <<= 1) s += __shfl_xor(s, k, 64);
    z[i] = s + prm[MLP_B2];
  }
}

// ---------------------------------------------------------------------------
// inference over the plain frontier: keep[v] = P(move) > tau
// ---------------------------------------------------------------------------
template <class T>
__global__ __launch_bounds__(256) void k_mlp_filter(DevGraph<T> g, DevState<T> st, const int32_t* __restrict__ front,
                                                     int nfront, int it, const float* __restrict__ prm,
                                                     const short* __restrict__ w1b, float tau, int dmin,
                                                     uint8_t* __restrict__ keep) {
  const int lane = threadIdx.x & 63;
  const int nw = gridDim.x * (blockDim.x >> 6);
  for (int base = (blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) * 16; base < nfront; base += nw * 16) {
    const int r = base + (lane & 15);
    int v = -1;
    int64_t d = 0;
    if (r < nfront) {
      v = front[r];
      d = g.end[v] - g.off[v];
    }
    // rows shorter than dmin are kept without a prediction: a wave whose 16
    // rows are all short (the frontier is grouped by degree bin, so most
    // waves of the short-row bins are) skips the features and the MFMAs
    if (!__any(r < nfront && d >= dmin)) {
      if (lane < 16 && r < nfront) keep[v] = 1;
      continue;
    }
    float x[4] = {0.f, 0.f, 0.f, 0.f};
    if (r < nfront) {
      float f[MLP_IN];
      vertex_features<T>(g, st, v, it, f);
#pragma unroll
      for (int j = 0; j < 4; j++) {
        const int k = 4 * (lane >> 4) + j;
        x[j] = k < MLP_IN ? f[k] : 0.f;
      }
    }
    f32x4 acc[4];
    mlp_hidden(x, w1b, acc);
    float z[4];
    mlp_logits(acc, prm, z);
    if ((lane & 15) == 0) {
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int rr = base + 4 * (lane >> 4) + i;
        if (rr < nfront) {
          const int v = front[rr];
          keep[v] = (sigmoidf(z[i]) > tau || g.end[v] - g.off[v] < dmin) ? 1 : 0;
        }
      }
    }
  }
}

// features of the frontier vertices (training logs)
template <class T>
__global__ void k_features(DevGraph<T> g, DevState<T> st, const int32_t* __restrict__ front, int nfront, int it,
                           float* __restrict__ out) {
  for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < nfront; r += gridDim.x * blockDim.x) {
    float f[MLP_IN];
    vertex_features<T>(g, st, front[r], it, f);
    for (int k = 0; k < MLP_IN; k++) out[(int64_t)r * MLP_IN + k] = f[k];
  }
}

__global__ void k_labels(const int32_t* __restrict__ front, int nfront, const int32_t* __restrict__ moved_it, int it,
                         uint8_t* __restrict__ lab) {
  for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < nfront; r += gridDim.x * blockDim.x)
    lab[r] = moved_it[front[r]] == it ? 1 : 0;
}

// ---------------------------------------------------------------------------
// probabilities for a feature matrix (tests)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_mlp_forward(const float* __restrict__ X, int n, const float* __restrict__ prm,
                                                      const short* __restrict__ w1b, float* __restrict__ p) {
  const int lane = threadIdx.x & 63;
  const int nw = gridDim.x * (blockDim.x >> 6);
  for (int base = (blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) * 16; base < n; base += nw * 16) {
    const int r = base + (lane & 15);
    float x[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int k = 4 * (lane >> 4) + j;
      if (r < n && k < MLP_IN) x[j] = X[(int64_t)r * MLP_IN + k];
    }
    f32x4 acc[4];
    mlp_hidden(x, w1b, acc);
    float z[4];
    mlp_logits(acc, prm, z);
    if ((lane & 15) == 0)
      for (int i = 0; i < 4; i++) {
        const int rr = base + 4 * (lane >> 4) + i;
        if (rr < n) p[rr] = sigmoidf(z[i]);
      }
  }
}

// ---------------------------------------------------------------------------
// gradients of the weighted binary cross-entropy over samples idx[0..n)
//   loss = sum_i c_i BCE(sigmoid(z_i), y_i) / sum_i c_i,  c_i = pos_w if y_i else 1
// grad[] (MLP_NPARAM floats) and loss[0] must be zero on entry; wsum = sum c_i.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_mlp_grad(const float* __restrict__ X, const uint8_t* __restrict__ y,
                                                   const int32_t* __restrict__ idx, int n, const float* __restrict__ prm,
                                                   const short* __restrict__ w1b, float pos_w, float inv_wsum,
                                                   float* __restrict__ grad, float* __restrict__ loss) {
  const int lane = threadIdx.x & 63;
  const int nw = gridDim.x * (blockDim.x >> 6);
  const float* b1 = prm + MLP_B1;
  const float* w2 = prm + MLP_W2;
  f32x4 gW1[4];
  for (int t = 0; t < 4; t++) gW1[t] = f32x4{0.f, 0.f, 0.f, 0.f};
  float gb1[4] = {0.f, 0.f, 0.f, 0.f}, gw2[4] = {0.f, 0.f, 0.f, 0.f};
  float gb2 = 0.f, lsum = 0.f;
  for (int base = (blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) * 16; base < n; base += nw * 16) {
    // forward operand: feature 4(l/16)+j of sample l%16
    const int r = base + (lane & 15);
    const int sr = r < n ? idx[r] : -1;
    float x[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int k = 4 * (lane >> 4) + j;
      if (sr >= 0 && k < MLP_IN) x[j] = X[(int64_t)sr * MLP_IN + k];
    }
    f32x4 acc[4];
    mlp_hidden(x, w1b, acc);
    float z[4];
    mlp_logits(acc, prm, z);
    // per-sample output gradient for rows 4(l/16)+i
    float dz[4];
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int rr = base + 4 * (lane >> 4) + i;
      dz[i] = 0.f;
      if (rr < n) {
        const int s = idx[rr];
        const float yy = y[s] ? 1.f : 0.f;
        const float c = y[s] ? pos_w : 1.f;
        const float p = sigmoidf(z[i]);
        dz[i] = c * (p - yy) * inv_wsum;
        // BCE with logits, numerically stable
        const float zz = z[i];
        const float l = (zz > 0.f ? zz : 0.f) - zz * yy + __logf(1.f + __expf(-fabsf(zz)));
        if ((lane & 15) == 0) lsum += c * l * inv_wsum;
        if ((lane & 15) == 0) gb2 += dz[i];
      }
    }
    // A operand of dW1 = X^T dH: feature l%16 of samples 4(l/16)+i
    bf16x4 xa;
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int rr = base + 4 * (lane >> 4) + i;
      const int k = lane & 15;
      float v = 0.f;
      if (rr < n && k < MLP_IN) v = X[(int64_t)idx[rr] * MLP_IN + k];
      xa[i] = to_bf16(v);
    }
#pragma unroll
    for (int t = 0; t < 4; t++) {
      const int h = 16 * t + (lane & 15);
      bf16x4 dh;
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const float a = acc[t][i] + b1[h];
        const float act = a > 0.f ? a : 0.f;
        gw2[t] += dz[i] * act;
        const float g = a > 0.f ? dz[i] * w2[h] : 0.f;
        gb1[t] += g;
        dh[i] = to_bf16(g);
      }
      gW1[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(xa, dh, gW1[t], 0, 0, 0);
    }
  }
  // reduce over the 4 row groups (lanes l, l^16, l^32, l^48 share hidden l%16)
#pragma unroll
  for (int t = 0; t < 4; t++) {
    gb1[t] += __shfl_xor(gb1[t], 16, 64);
    gb1[t] += __shfl_xor(gb1[t], 32, 64);
    gw2[t] += __shfl_xor(gw2[t], 16, 64);
    gw2[t] += __shfl_xor(gw2[t], 32, 64);
  }
  gb2 = wave_sum(gb2);
  lsum = wave_sum(lsum);
#pragma unroll
  for (int t = 0; t < 4; t++) {
    const int h = 16 * t + (lane & 15);
    // gW1[t][i] = dW1[feature 4(l/16)+i][hidden h]
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int k = 4 * (lane >> 4) + i;
      if (k < MLP_IN) atomicAdd(&grad[MLP_W1 + k * MLP_H + h], gW1[t][i]);
    }
    if (lane < 16) {
      atomicAdd(&grad[MLP_B1 + h], gb1[t]);
      atomicAdd(&grad[MLP_W2 + h], gw2[t]);
    }
  }
  if (lane == 0) {
    atomicAdd(&grad[MLP_B2], gb2);
    atomicAdd(loss, lsum);
  }
}

// Adam on the fp32 master parameters; refreshes the BF16 copy of W1.
__global__ void k_adam(float* __restrict__ prm, float* __restrict__ m, float* __restrict__ v,
                       const float* __restrict__ grad, float lr, float b1, float b2, float eps, int step,
                       short* __restrict__ w1b) {
  const float c1 = 1.f - __powf(b1, (float)step), c2 = 1.f - __powf(b2, (float)step);
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < MLP_NPARAM; i += gridDim.x * blockDim.x) {
    const bool w1_pad = i < MLP_B1 && (i / MLP_H) >= MLP_IN;  // padded feature rows stay zero
    if (!w1_pad) {
      const float g = grad[i];
      const float mm = b1 * m[i] + (1.f - b1) * g;
      const float vv = b2 * v[i] + (1.f - b2) * g * g;
      m[i] = mm;
      v[i] = vv;
      prm[i] -= lr * (mm / c1) / (__fsqrt_rn(vv / c2) + eps);
    }
    if (i < MLP_B1) w1b[i] = to_bf16(prm[i]);
  }
}

__global__ void k_w1_bf16(const float* __restrict__ prm, short* __restrict__ w1b) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < MLP_B1; i += gridDim.x * blockDim.x)
    w1b[i] = to_bf16(prm[i]);
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
template <class T>
void launch_mlp_filter(const DevGraph<T>& g, DevState<T>& st, const int32_t* front, int nfront, int it,
                       const float* prm, const short* w1b, float tau, int dmin, uint8_t* keep, hipStream_t s) {
  if (nfront <= 0) return;
  const int waves = (nfront + 15) / 16;
  k_mlp_filter<T><<<std::min((waves + 3) / 4, 8192), 256, 0, s>>>(g, st, front, nfront, it, prm, w1b, tau, dmin,
                                                                   keep);
}

template <class T>
void launch_features(const DevGraph<T>& g, DevState<T>& st, const int32_t* front, int nfront, int it, float* out,
                     hipStream_t s) {
  if (nfront <= 0) return;
  k_features<T><<<std::min((nfront + 255) / 256, 8192), 256, 0, s>>>(g, st, front, nfront, it, out);
}

void launch_labels(const int32_t* front, int nfront, const int32_t* moved_it, int it, uint8_t* lab, hipStream_t s) {
  if (nfront <= 0) return;
  k_labels<<<std::min((nfront + 255) / 256, 8192), 256, 0, s>>>(front, nfront, moved_it, it, lab);
}

void launch_mlp_forward(const float* X, int n, const float* prm, const short* w1b, float* p, hipStream_t s) {
  if (n <= 0) return;
  const int waves = (n + 15) / 16;
  k_mlp_forward<<<std::min((waves + 3) / 4, 8192), 256, 0, s>>>(X, n, prm, w1b, p);
}

void launch_mlp_grad(const float* X, const uint8_t* y, const int32_t* idx, int n, const float* prm, const short* w1b,
                     float pos_w, float inv_wsum, float* grad, float* loss, hipStream_t s) {
  if (n <= 0) return;
  const int waves = (n + 15) / 16;
  k_mlp_grad<<<std::min((waves + 3) / 4, 1024), 256, 0, s>>>(X, y, idx, n, prm, w1b, pos_w, inv_wsum, grad, loss);
}

void launch_adam(float* prm, float* m, float* v, const float* grad, float lr, int step, short* w1b, hipStream_t s) {
  k_adam<<<(MLP_NPARAM + 255) / 256, 256, 0, s>>>(prm, m, v, grad, lr, 0.9f, 0.999f, 1e-8f, step, w1b);
}

void launch_w1_bf16(const float* prm, short* w1b, hipStream_t s) {
  k_w1_bf16<<<(MLP_B1 + 255) / 256, 256, 0, s>>>(prm, w1b);
}

template void launch_mlp_filter<IntW>(const DevGraph<IntW>&, DevState<IntW>&, const int32_t*, int, int, const float*,
                                      const short*, float, int, uint8_t*, hipStream_t);
template void launch_mlp_filter<WideW>(const DevGraph<WideW>&, DevState<WideW>&, const int32_t*, int, int, const float*,
                                      const short*, float, int, uint8_t*, hipStream_t);
template void launch_features<IntW>(const DevGraph<IntW>&, DevState<IntW>&, const int32_t*, int, int, float*,
                                    hipStream_t);
template void launch_features<WideW>(const DevGraph<WideW>&, DevState<WideW>&, const int32_t*, int, int, float*,
                                    hipStream_t);

}  // namespace gala
