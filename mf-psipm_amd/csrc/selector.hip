// Learned column selector: a per-column MLP 8 -> 64 -> 64 -> 1 (ReLU) that
// scores candidate columns of the active-column preconditioner from their
// interior-point trajectory.  The reference selects columns by the leverage
// rule alone (psipm_stable_preconditioner.py:293-371); this network is an
// addition whose output replaces the leverage scores in the same retention /
// fill procedure, with the rule as the fallback.
//
//   * features (k_select_features): per candidate column j
//       log10 theta_j at the current and the two previous factorisations,
//       log10 x_j, log10 z_j, log10 ||a_j||_2, log10 mu, iteration index;
//     normalised by the model's (mean, 1/std).
//   * inference (k_mlp_infer): BF16 MFMA (16x16x16) for both hidden layers
//     with FP32 accumulation; a wave scores 16 columns per pass with the
//     weights held in registers, the hidden tile moves through LDS between
//     the layers.
//   * training (k_mlp_grad, k_mlp_reduce, k_mlp_adam): FP32 forward and
//     backward per sample, per-block gradient partials summed in a fixed
//     order (bitwise reproducible), Adam.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <random>
#include <vector>

#include "ctx.hpp"

namespace psg {

constexpr int kIn = 8, kHid = 64;
constexpr int kOffW1 = 0;
constexpr int kOffB1 = kOffW1 + kIn * kHid;
constexpr int kOffW2 = kOffB1 + kHid;
constexpr int kOffB2 = kOffW2 + kHid * kHid;
constexpr int kOffW3 = kOffB2 + kHid;
constexpr int kOffB3 = kOffW3 + kHid;
constexpr int kNParams = kOffB3 + 1;  // 4801
constexpr int kTB = 128;               // samples per training block
constexpr int kOwn = (kNParams + kTB - 1) / kTB;  // gradient entries per thread

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// float -> bf16 bits, round to nearest even (finite inputs)
__device__ __host__ inline unsigned short to_bf16(float f) {
  unsigned int u;
  memcpy(&u, &f, 4);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}

}  // namespace psg

struct psg_mlp {
  hipStream_t stream = nullptr;
  psg::DevBuf<float> params, m1, m2, grad, part, norm;
  psg::DevBuf<unsigned short> wb;  // bf16 W1 padded to 16 x 64, then W2 64 x 64 (row-major [k][j])
  psg::DevBuf<float> X, y, out;
  psg::DevBuf<int> idx;
  int step = 0;
  bool have_params = false;
};

namespace psg {

// ---------------------------------------------------------------------------
// features
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock)
k_select_features(int ncand, const int* __restrict__ cand, const int* __restrict__ cptr,
                  const double* __restrict__ cval, const double* __restrict__ th0,
                  const double* __restrict__ th1, const double* __restrict__ th2,
                  const double* __restrict__ x, const double* __restrict__ z, float log_mu, float iter,
                  float* __restrict__ out) {
  const double tiny = 1e-300;
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < ncand; k += (int64_t)gridDim.x * blockDim.x) {
    const int j = cand[k];
    double nrm = 0.0;
    for (int p = cptr[j]; p < cptr[j + 1]; ++p) nrm += cval[p] * cval[p];
    float f[kIn];
    f[0] = (float)log10(fmax(th0[j], tiny));
    f[1] = (float)log10(fmax(th1[j], tiny));
    f[2] = (float)log10(fmax(th2[j], tiny));
    f[3] = (float)log10(fmax(x[j], tiny));
    f[4] = (float)log10(fmax(z[j], tiny));
    f[5] = (float)(0.5 * log10(fmax(nrm, tiny)));
    f[6] = log_mu;
    f[7] = iter;
#pragma unroll
    for (int i = 0; i < kIn; ++i) out[k * kIn + i] = f[i];
  }
}

// ---------------------------------------------------------------------------
// inference: BF16 MFMA, FP32 accumulation
// ---------------------------------------------------------------------------
__global__ void k_mlp_pack(const float* __restrict__ P, unsigned short* __restrict__ wb) {
  // W1 padded: rows k < 8 from W1, rows 8..15 zero; then W2
  for (int e = threadIdx.x + blockIdx.x * blockDim.x; e < 16 * kHid + kHid * kHid; e += gridDim.x * blockDim.x) {
    float v;
    if (e < 16 * kHid) {
      const int k = e / kHid, j = e % kHid;
      v = k < kIn ? P[kOffW1 + k * kHid + j] : 0.f;
    } else {
      v = P[kOffW2 + (e - 16 * kHid)];
    }
    wb[e] = to_bf16(v);
  }
}

// v_mfma_f32_16x16x16_bf16 operand layout: A(i, k) in lane i + 16 (k / 4),
// element k % 4; B(k, j) in lane j + 16 (k / 4), element k % 4; result
// D(4 (lane / 16) + v, lane % 16) in element v.
__global__ void __launch_bounds__(kBlock)
k_mlp_infer(int B, const float* __restrict__ X, const float* __restrict__ norm, const float* __restrict__ P,
            const unsigned short* __restrict__ wb, float* __restrict__ out) {
  __shared__ float sH[kBlock / kWave][16][kHid + 4];
  const int lane = threadIdx.x & (kWave - 1), wv = threadIdx.x / kWave;
  const int li = lane & 15, lg = lane >> 4;
  // weights as B operands, held for the whole launch
  bf16x4 w1[4], w2[4][4];
#pragma unroll
  for (int nt = 0; nt < 4; ++nt)
#pragma unroll
    for (int t = 0; t < 4; ++t) w1[nt][t] = (short)wb[(4 * lg + t) * kHid + nt * 16 + li];
#pragma unroll
  for (int kt = 0; kt < 4; ++kt)
#pragma unroll
    for (int nt = 0; nt < 4; ++nt)
#pragma unroll
      for (int t = 0; t < 4; ++t)
        w2[kt][nt][t] = (short)wb[16 * kHid + (kt * 16 + 4 * lg + t) * kHid + nt * 16 + li];
  float b1[4], b2[4], w3[4];
#pragma unroll
  for (int nt = 0; nt < 4; ++nt) {
    b1[nt] = P[kOffB1 + nt * 16 + li];
    b2[nt] = P[kOffB2 + nt * 16 + li];
    w3[nt] = P[kOffW3 + nt * 16 + li];
  }
  const float b3 = P[kOffB3];
  float mean[4], istd[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const int k = 4 * lg + t;
    mean[t] = k < kIn ? norm[k] : 0.f;
    istd[t] = k < kIn ? norm[kIn + k] : 0.f;
  }
  const int waves = gridDim.x * (kBlock / kWave);
  for (int64_t tile = (int64_t)blockIdx.x * (kBlock / kWave) + wv; tile * 16 < B; tile += waves) {
    const int64_t r0 = tile * 16;
    // layer 1 A operand: features of sample r0 + li, k = 4 lg + t (k < 8)
    bf16x4 a;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int k = 4 * lg + t;
      const int64_t r = r0 + li;
      float v = 0.f;
      if (k < kIn && r < B) v = (X[r * kIn + k] - mean[t]) * istd[t];
      a[t] = (short)to_bf16(v);
    }
    f32x4 h[4];
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) {
      f32x4 c = {b1[nt], b1[nt], b1[nt], b1[nt]};
      h[nt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, w1[nt], c, 0, 0, 0);
    }
    // relu, hidden tile to LDS as (sample, unit)
#pragma unroll
    for (int nt = 0; nt < 4; ++nt)
#pragma unroll
      for (int v = 0; v < 4; ++v) sH[wv][4 * lg + v][nt * 16 + li] = fmaxf(h[nt][v], 0.f);
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_s_waitcnt(0);
    f32x4 g[4];
#pragma unroll
    for (int nt = 0; nt < 4; ++nt) g[nt] = f32x4{b2[nt], b2[nt], b2[nt], b2[nt]};
#pragma unroll
    for (int kt = 0; kt < 4; ++kt) {
      bf16x4 a2;
#pragma unroll
      for (int t = 0; t < 4; ++t) a2[t] = (short)to_bf16(sH[wv][li][kt * 16 + 4 * lg + t]);
#pragma unroll
      for (int nt = 0; nt < 4; ++nt) g[nt] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a2, w2[kt][nt], g[nt], 0, 0, 0);
    }
    __builtin_amdgcn_wave_barrier();  // LDS reads of this tile done before the next tile's writes
    // output layer: row 4 lg + v, partial over units nt*16 + li
    float s[4];
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      float acc = 0.f;
#pragma unroll
      for (int nt = 0; nt < 4; ++nt) acc += fmaxf(g[nt][v], 0.f) * w3[nt];
      s[v] = acc;
    }
#pragma unroll
    for (int v = 0; v < 4; ++v)
#pragma unroll
      for (int o = 8; o > 0; o >>= 1) s[v] += __shfl_xor(s[v], o, 64);
    if (li == 0) {
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int64_t r = r0 + 4 * lg + v;
        if (r < B) out[r] = s[v] + b3;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// training: FP32 forward + backward, deterministic gradient sums
// ---------------------------------------------------------------------------
// One block takes chunks of kTB samples (thread = sample for the per-sample
// passes); each thread owns gradient entries e = tid + kTB q and sums them
// over the block's samples.  part[blockIdx.x * (kNParams + 1) + e] receives
// the block's sums; entry kNParams holds the block's loss sum.
__global__ void __launch_bounds__(kTB)
k_mlp_grad(int B, const float* __restrict__ X, const float* __restrict__ Y, const int* __restrict__ idx,
           const float* __restrict__ norm, const float* __restrict__ P, float pos_weight, float inv_total,
           float* __restrict__ part) {
  __shared__ float sP[kNParams];
  __shared__ float sX[kTB][kIn + 1];
  __shared__ float sH1[kTB][kHid + 1];
  __shared__ float sD2[kTB][kHid + 1];
  __shared__ float sT[kTB][kHid + 1];  // h2, later d1
  __shared__ float sDz[kTB];
  __shared__ float sOne[1];
  __shared__ float sLoss[kTB];
  const int tid = threadIdx.x;
  for (int e = tid; e < kNParams; e += kTB) sP[e] = P[e];
  if (tid == 0) sOne[0] = 1.f;
  float acc[kOwn];
#pragma unroll
  for (int q = 0; q < kOwn; ++q) acc[q] = 0.f;
  float loss_acc = 0.f;
  __syncthreads();
  for (int64_t c0 = (int64_t)blockIdx.x * kTB; c0 < B; c0 += (int64_t)gridDim.x * kTB) {
    const int64_t sidx = c0 + tid;
    const bool valid = sidx < B;
    // --- per-sample forward and the output-layer backward
    float yv = 0.f;
    if (valid) {
      const int64_t row = idx ? (int64_t)idx[sidx] : sidx;
#pragma unroll
      for (int k = 0; k < kIn; ++k) sX[tid][k] = (X[row * kIn + k] - norm[k]) * norm[kIn + k];
      yv = Y[row];
    } else {
#pragma unroll
      for (int k = 0; k < kIn; ++k) sX[tid][k] = 0.f;
    }
    for (int j = 0; j < kHid; ++j) {
      float a = sP[kOffB1 + j];
#pragma unroll
      for (int k = 0; k < kIn; ++k) a += sX[tid][k] * sP[kOffW1 + k * kHid + j];
      sH1[tid][j] = fmaxf(a, 0.f);
    }
    float zz = sP[kOffB3];
    for (int j = 0; j < kHid; ++j) {
      float a = sP[kOffB2 + j];
      for (int i = 0; i < kHid; ++i) a += sH1[tid][i] * sP[kOffW2 + i * kHid + j];
      const float h2 = fmaxf(a, 0.f);
      sT[tid][j] = h2;
      zz += h2 * sP[kOffW3 + j];
    }
    // weighted BCE with logits; softplus(t) = max(t, 0) + log1p(exp(-|t|))
    const float spn = fmaxf(-zz, 0.f) + log1pf(expf(-fabsf(zz)));  // softplus(-z)
    const float spp = spn + zz;                                      // softplus(z)
    const float sig = 1.f / (1.f + expf(-zz));
    float dz = 0.f;
    if (valid) {
      loss_acc += pos_weight * yv * spn + (1.f - yv) * spp;
      dz = ((1.f - yv) * sig - pos_weight * yv * (1.f - sig)) * inv_total;
    }
    sDz[tid] = dz;
    for (int j = 0; j < kHid; ++j) sD2[tid][j] = sT[tid][j] > 0.f ? dz * sP[kOffW3 + j] : 0.f;
    __syncthreads();
    // --- gradients of W2, b2, w3, b3 over the chunk
#pragma unroll
    for (int q = 0; q < kOwn; ++q) {
      const int e = tid + kTB * q;
      if (e < kOffW2 || e >= kNParams) continue;
      const float* L;
      const float* R;
      int ls, rs;
      if (e < kOffB2) { const int i = (e - kOffW2) / kHid, j = (e - kOffW2) % kHid; L = &sH1[0][i]; ls = kHid + 1; R = &sD2[0][j]; rs = kHid + 1; }
      else if (e < kOffW3) { const int j = e - kOffB2; L = sOne; ls = 0; R = &sD2[0][j]; rs = kHid + 1; }
      else if (e < kOffB3) { const int j = e - kOffW3; L = &sT[0][j]; ls = kHid + 1; R = sDz; rs = 1; }
      else { L = sOne; ls = 0; R = sDz; rs = 1; }
      float sacc = 0.f;
      for (int s = 0; s < kTB; ++s) sacc += L[s * ls] * R[s * rs];
      acc[q] += sacc;
    }
    __syncthreads();
    // --- d1 = (d2 W2^T) o [h1 > 0]
    for (int i = 0; i < kHid; ++i) {
      float a = 0.f;
      if (sH1[tid][i] > 0.f)
        for (int j = 0; j < kHid; ++j) a += sD2[tid][j] * sP[kOffW2 + i * kHid + j];
      sT[tid][i] = a;
    }
    __syncthreads();
    // --- gradients of W1, b1
#pragma unroll
    for (int q = 0; q < kOwn; ++q) {
      const int e = tid + kTB * q;
      if (e >= kOffW2) continue;
      const float* L;
      int ls;
      int j;
      if (e < kOffB1) { const int k = e / kHid; j = e % kHid; L = &sX[0][k]; ls = kIn + 1; }
      else { j = e - kOffB1; L = sOne; ls = 0; }
      float sacc = 0.f;
      for (int s = 0; s < kTB; ++s) sacc += L[s * ls] * sT[s][j];
      acc[q] += sacc;
    }
    __syncthreads();
  }
  float* dst = part + (size_t)blockIdx.x * (kNParams + 1);
#pragma unroll
  for (int q = 0; q < kOwn; ++q) {
    const int e = tid + kTB * q;
    if (e < kNParams) dst[e] = acc[q];
  }
  sLoss[tid] = loss_acc;
  __syncthreads();
  if (tid == 0) {
    float l = 0.f;
    for (int t = 0; t < kTB; ++t) l += sLoss[t];
    dst[kNParams] = l;
  }
}

// grad[e] = sum over blocks (fixed order); grad[kNParams] = loss sum
__global__ void k_mlp_reduce(int nb, const float* __restrict__ part, float* __restrict__ grad) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e > kNParams) return;
  float s = 0.f;
  for (int b = 0; b < nb; ++b) s += part[(size_t)b * (kNParams + 1) + e];
  grad[e] = s;
}

// Adam (Kingma & Ba 2015, Algorithm 1)
__global__ void k_mlp_adam(float* __restrict__ P, float* __restrict__ m1, float* __restrict__ m2,
                           const float* __restrict__ g, float lr, float beta1, float beta2, float eps,
                           float c1, float c2) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= kNParams) return;
  const float ge = g[e];
  const float a = beta1 * m1[e] + (1.f - beta1) * ge;
  const float b = beta2 * m2[e] + (1.f - beta2) * ge * ge;
  m1[e] = a;
  m2[e] = b;
  P[e] -= lr * (a / c1) / (sqrtf(b / c2) + eps);
}

static int grad_grid(int64_t B) { return (int)std::max<int64_t>(1, std::min<int64_t>(256, (B + kTB - 1) / kTB)); }

static void mlp_pack(psg_mlp* net) {
  net->wb.grow(16 * kHid + kHid * kHid);
  hipLaunchKernelGGL(k_mlp_pack, dim3(20), dim3(kBlock), 0, net->stream, net->params.p, net->wb.p);
  HIP_LAUNCH_CHECK();
}

static void mlp_infer(psg_mlp* net, int64_t B, const float* Xd, float* outd, hipStream_t s) {
  if (B <= 0) return;
  const int64_t tiles = (B + 15) / 16;
  const int grid = (int)std::max<int64_t>(1, std::min<int64_t>(4096, (tiles + 3) / 4));
  hipLaunchKernelGGL(k_mlp_infer, dim3(grid), dim3(kBlock), 0, s, (int)B, Xd, net->norm.p, net->params.p,
                     net->wb.p, outd);
  HIP_LAUNCH_CHECK();
}

// loss sum and gradient (mean over the batch) into net->grad
static void mlp_grad(psg_mlp* net, int64_t B, const float* Xd, const float* Yd, const int* idxd, float pos_weight) {
  const int nb = grad_grid(B);
  net->part.grow((size_t)nb * (kNParams + 1));
  net->grad.grow(kNParams + 1);
  hipLaunchKernelGGL(k_mlp_grad, dim3(nb), dim3(kTB), 0, net->stream, (int)B, Xd, Yd, idxd, net->norm.p,
                     net->params.p, pos_weight, 1.f / (float)B, net->part.p);
  HIP_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_mlp_reduce, dim3((kNParams + 1 + kBlock - 1) / kBlock), dim3(kBlock), 0, net->stream, nb,
                     net->part.p, net->grad.p);
  HIP_LAUNCH_CHECK();
}

static void need_params(const psg_mlp* net) {
  if (!net->have_params) throw Error(PSG_ERR_STATE, "selector parameters not set (psg_mlp_set_params)");
}

}  // namespace psg

using namespace psg;

extern "C" {

int psg_mlp_create(psg_mlp** out) {
  return guarded([&]() -> int {
    if (!out) throw Error(PSG_ERR_ARGUMENT, "psg_mlp_create: null output");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) throw Error(PSG_ERR_DEVICE, "psg_mlp_create: no GPU visible");
    auto* net = new psg_mlp;
    try {
      HIP_CHECK(hipStreamCreateWithFlags(&net->stream, hipStreamNonBlocking));
      net->params.alloc(kNParams);
      net->m1.alloc(kNParams);
      net->m2.alloc(kNParams);
      net->norm.alloc(2 * kIn);
    } catch (...) {
      psg_mlp_destroy(net);
      throw;
    }
    *out = net;
    return PSG_OK;
  });
}

int psg_mlp_destroy(psg_mlp* net) {
  if (!net) return PSG_OK;
  if (net->stream) (void)hipStreamSynchronize(net->stream);
  net->params.release();
  net->m1.release();
  net->m2.release();
  net->grad.release();
  net->part.release();
  net->norm.release();
  net->wb.release();
  net->X.release();
  net->y.release();
  net->out.release();
  net->idx.release();
  if (net->stream) (void)hipStreamDestroy(net->stream);
  delete net;
  return PSG_OK;
}

int psg_mlp_num_params(void) { return kNParams; }

int psg_mlp_set_params(psg_mlp* net, const float* params, const float* norm) {
  return guarded([&]() -> int {
    if (!net || !params || !norm) throw Error(PSG_ERR_ARGUMENT, "psg_mlp_set_params: null argument");
    for (int i = 0; i < kNParams; ++i)
      if (!std::isfinite(params[i])) throw Error(PSG_ERR_ARGUMENT, "psg_mlp_set_params: non-finite parameter");
    net->params.upload(params, kNParams, net->stream);
    net->norm.upload(norm, 2 * kIn, net->stream);
    std::vector<float> zero(kNParams, 0.f);
    net->m1.upload(zero, net->stream);
    net->m2.upload(zero, net->stream);
    net->step = 0;
    mlp_pack(net);
    HIP_CHECK(hipStreamSynchronize(net->stream));
    net->have_params = true;
    return PSG_OK;
  });
}

int psg_mlp_get_params(psg_mlp* net, float* params) {
  return guarded([&]() -> int {
    if (!net || !params) throw Error(PSG_ERR_ARGUMENT, "psg_mlp_get_params: null argument");
    need_params(net);
    net->params.download(params, kNParams, net->stream);
    HIP_CHECK(hipStreamSynchronize(net->stream));
    return PSG_OK;
  });
}

int psg_mlp_forward(psg_mlp* net, const float* X, int64_t B, float* logits) {
  return guarded([&]() -> int {
    if (!net || B < 0 || (B && (!X || !logits))) throw Error(PSG_ERR_ARGUMENT, "psg_mlp_forward: invalid arguments");
    need_params(net);
    if (B == 0) return PSG_OK;
    if (B >= (1LL << 31) / kIn) throw Error(PSG_ERR_ARGUMENT, "psg_mlp_forward: batch too large");
    net->X.upload(X, (size_t)B * kIn, net->stream);
    net->out.grow(B);
    mlp_infer(net, B, net->X.p, net->out.p, net->stream);
    net->out.download(logits, B, net->stream);
    HIP_CHECK(hipStreamSynchronize(net->stream));
    return PSG_OK;
  });
}

int psg_mlp_loss_grad(psg_mlp* net, const float* X, const float* y, int64_t B, float pos_weight, float* loss,
                      float* grad) {
  return guarded([&]() -> int {
    if (!net || B < 1 || !X || !y || !loss || !grad) throw Error(PSG_ERR_ARGUMENT, "psg_mlp_loss_grad: invalid arguments");
    need_params(net);
    if (B >= (1LL << 31) / kIn) throw Error(PSG_ERR_ARGUMENT, "psg_mlp_loss_grad: batch too large");
    net->X.upload(X, (size_t)B * kIn, net->stream);
    net->y.upload(y, B, net->stream);
    mlp_grad(net, B, net->X.p, net->y.p, nullptr, pos_weight);
    std::vector<float> g(kNParams + 1);
    net->grad.download(g.data(), kNParams + 1, net->stream);
    HIP_CHECK(hipStreamSynchronize(net->stream));
    std::copy(g.begin(), g.begin() + kNParams, grad);
    *loss = g[kNParams] / (float)B;
    return PSG_OK;
  });
}

int psg_mlp_train(psg_mlp* net, const float* X, const float* y, int64_t N, int32_t batch, int32_t steps,
                  float lr, float pos_weight, uint64_t seed, float* losses) {
  return guarded([&]() -> int {
    if (!net || N < 1 || !X || !y || batch < 1 || steps < 0 || !(lr > 0.f))
      throw Error(PSG_ERR_ARGUMENT, "psg_mlp_train: invalid arguments");
    need_params(net);
    if (N >= (1LL << 31) / kIn) throw Error(PSG_ERR_ARGUMENT, "psg_mlp_train: dataset too large");
    hipStream_t s = net->stream;
    net->X.upload(X, (size_t)N * kIn, s);
    net->y.upload(y, N, s);
    const int64_t bsz = std::min<int64_t>(batch, N);
    // epoch permutations drawn on the host, uploaded once per epoch
    std::mt19937_64 rng(seed);
    std::vector<int> perm(N);
    std::iota(perm.begin(), perm.end(), 0);
    int64_t pos = N;
    std::vector<float> hloss(steps > 0 ? steps : 1);
    std::vector<float> gl(1);
    const float beta1 = 0.9f, beta2 = 0.999f, eps = 1e-8f;
    net->idx.grow(N);
    for (int st = 0; st < steps; ++st) {
      if (pos + bsz > N) {
        std::shuffle(perm.begin(), perm.end(), rng);
        net->idx.upload(perm, s);
        pos = 0;
      }
      mlp_grad(net, bsz, net->X.p, net->y.p, net->idx.p + pos, pos_weight);
      pos += bsz;
      net->step += 1;
      const float c1 = 1.f - std::pow(beta1, (float)net->step);
      const float c2 = 1.f - std::pow(beta2, (float)net->step);
      hipLaunchKernelGGL(k_mlp_adam, dim3((kNParams + kBlock - 1) / kBlock), dim3(kBlock), 0, s, net->params.p,
                         net->m1.p, net->m2.p, net->grad.p, lr, beta1, beta2, eps, c1, c2);
      HIP_LAUNCH_CHECK();
      if (losses) {
        HIP_CHECK(hipMemcpyAsync(&hloss[st], net->grad.p + kNParams, sizeof(float), hipMemcpyDeviceToHost, s));
      }
    }
    mlp_pack(net);
    HIP_CHECK(hipStreamSynchronize(s));
    if (losses)
      for (int st = 0; st < steps; ++st) losses[st] = hloss[st] / (float)bsz;
    return PSG_OK;
  });
}

// Features of the candidate columns from the context's weight history and
// the supplied iterate (x, z: n doubles), optionally downloaded, and (with a
// network) their logits.
int psg_select_features(psg_ctx* c, const int32_t* cand, int64_t ncand, const double* x, const double* z,
                        double mu, double iter, float* features) {
  return guarded([&]() -> int {
    if (!c || ncand < 0 || (ncand && (!cand || !x || !z || !features)))
      throw Error(PSG_ERR_ARGUMENT, "psg_select_features: invalid arguments");
    if (!c->have_weights) throw Error(PSG_ERR_STATE, "weights not set (call psg_set_weights)");
    if (ncand == 0) return PSG_OK;
    for (int64_t k = 0; k < ncand; ++k)
      if (cand[k] < 0 || cand[k] >= c->n) throw Error(PSG_ERR_ARGUMENT, "psg_select_features: candidate out of range");
    hipStream_t s = c->stream;
    c->cand.upload(cand, ncand, s);
    c->sel_x.upload(x, c->n, s);
    c->sel_z.upload(z, c->n, s);
    c->sel_feat.grow((size_t)ncand * kIn);
    const double* h1 = c->nhist >= 2 ? c->theta_h1.p : c->theta.p;
    const double* h2 = c->nhist >= 3 ? c->theta_h2.p : h1;
    hipLaunchKernelGGL(k_select_features, dim3(grid_for(ncand, kBlock, 8192)), dim3(kBlock), 0, s, (int)ncand,
                       c->cand.p, c->cptr.p, c->cval.p, c->theta.p, h1, h2, c->sel_x.p, c->sel_z.p,
                       (float)std::log10(std::max(mu, 1e-300)), (float)iter, c->sel_feat.p);
    HIP_LAUNCH_CHECK();
    HIP_CHECK(hipMemcpyAsync(features, c->sel_feat.p, (size_t)ncand * kIn * sizeof(float), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    return PSG_OK;
  });
}

int psg_select_logits(psg_ctx* c, psg_mlp* net, const int32_t* cand, int64_t ncand, const double* x,
                      const double* z, double mu, double iter, float* logits) {
  return guarded([&]() -> int {
    if (!c || !net || ncand < 0 || (ncand && (!cand || !x || !z || !logits)))
      throw Error(PSG_ERR_ARGUMENT, "psg_select_logits: invalid arguments");
    if (!c->have_weights) throw Error(PSG_ERR_STATE, "weights not set (call psg_set_weights)");
    need_params(net);
    if (ncand == 0) return PSG_OK;
    for (int64_t k = 0; k < ncand; ++k)
      if (cand[k] < 0 || cand[k] >= c->n) throw Error(PSG_ERR_ARGUMENT, "psg_select_logits: candidate out of range");
    hipStream_t s = c->stream;
    // the network's parameters were written on its own stream
    HIP_CHECK(hipStreamSynchronize(net->stream));
    c->cand.upload(cand, ncand, s);
    c->sel_x.upload(x, c->n, s);
    c->sel_z.upload(z, c->n, s);
    c->sel_feat.grow((size_t)ncand * kIn);
    c->sel_out.grow(ncand);
    const double* h1 = c->nhist >= 2 ? c->theta_h1.p : c->theta.p;
    const double* h2 = c->nhist >= 3 ? c->theta_h2.p : h1;
    hipLaunchKernelGGL(k_select_features, dim3(grid_for(ncand, kBlock, 8192)), dim3(kBlock), 0, s, (int)ncand,
                       c->cand.p, c->cptr.p, c->cval.p, c->theta.p, h1, h2, c->sel_x.p, c->sel_z.p,
                       (float)std::log10(std::max(mu, 1e-300)), (float)iter, c->sel_feat.p);
    HIP_LAUNCH_CHECK();
    mlp_infer(net, ncand, c->sel_feat.p, c->sel_out.p, s);
    HIP_CHECK(hipMemcpyAsync(logits, c->sel_out.p, (size_t)ncand * sizeof(float), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    return PSG_OK;
  });
}

}  // extern "C"
