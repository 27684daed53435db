// Device-side utilities shared by the HIP translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "errors.hpp"

#define HIP_CHECK(expr)                                                                  \
  do {                                                                                   \
    hipError_t e_ = (expr);                                                              \
    if (e_ != hipSuccess)                                                                \
      throw psg::Error(PSG_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define HIP_LAUNCH_CHECK() HIP_CHECK(hipGetLastError())

namespace psg {

constexpr int kBlock = 256;       // threads per workgroup (4 wavefronts of 64)
constexpr int kWave = 64;
constexpr int kRedBlocks = 512;   // grid of reduction-producing vector kernels
constexpr int kSpmvBlocks = 2048; // grid cap of the row-gather kernels (partials finalized separately)
constexpr int kMaxRhs = 2;        // right-hand sides solved together (block PCG)

// RAII device buffer
template <class T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  ~DevBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  void alloc(size_t count) {
    if (count == n && p) return;
    release();
    if (count == 0) return;
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, count * sizeof(T));
    if (e != hipSuccess)
      throw Error(PSG_ERR_MEMORY, std::string("hipMalloc of ") + std::to_string(count * sizeof(T)) +
                                      " bytes failed: " + hipGetErrorString(e));
    p = static_cast<T*>(q);
    n = count;
  }
  void grow(size_t count) {
    if (count > n) alloc(count);
  }
  void upload(const T* h, size_t count, hipStream_t s) {
    grow(count);
    if (count) HIP_CHECK(hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, s));
  }
  void upload(const std::vector<T>& h, hipStream_t s) { upload(h.data(), h.size(), s); }
  void download(T* h, size_t count, hipStream_t s) const {
    if (count) HIP_CHECK(hipMemcpyAsync(h, p, count * sizeof(T), hipMemcpyDeviceToHost, s));
  }
};

inline int grid_for(int64_t work, int per_block, int cap) {
  int64_t g = (work + per_block - 1) / per_block;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return (int)g;
}

// ---- deterministic reductions ----------------------------------------------
__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Sum NP values over the block; every thread receives the totals.  The
// summation tree is fixed, so the result is bitwise reproducible.
template <int NP>
__device__ inline void block_sum(double (&v)[NP]) {
  __shared__ double sh[NP][kBlock / kWave];
  const int lane = threadIdx.x & (kWave - 1);
  const int wv = threadIdx.x / kWave;
#pragma unroll
  for (int k = 0; k < NP; ++k) {
    double s = wave_sum(v[k]);
    if (lane == 0) sh[k][wv] = s;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < NP; ++k) {
    double s = 0.0;
#pragma unroll
    for (int w = 0; w < kBlock / kWave; ++w) s += sh[k][w];
    v[k] = s;
  }
  __syncthreads();
}

// Every block of a consumer kernel re-reduces the producer's per-block
// partials (layout part[k * nb + b]) in the same fixed order, so all blocks
// hold identical totals without a grid-wide barrier or atomics.
template <int NP>
__device__ inline void reduce_partials(const double* __restrict__ part, int nb, double (&out)[NP]) {
#pragma unroll
  for (int k = 0; k < NP; ++k) out[k] = 0.0;
  for (int b = threadIdx.x; b < nb; b += blockDim.x) {
#pragma unroll
    for (int k = 0; k < NP; ++k) out[k] += part[(size_t)k * nb + b];
  }
  block_sum<NP>(out);
}

template <int NP>
__device__ inline void write_partials(double (&v)[NP], double* __restrict__ part) {
  block_sum<NP>(v);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NP; ++k) part[(size_t)k * gridDim.x + blockIdx.x] = v[k];
  }
}

// host-side fixed-order sum of per-block partials
inline double host_sum(const double* part, int nb) {
  double s = 0.0;
  for (int b = 0; b < nb; ++b) s += part[b];
  return s;
}

}  // namespace psg
