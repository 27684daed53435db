// Numeric supernodal Cholesky of the active-column matrix and its
// triangular solves (the preconditioner of the spectral stable-column PCG).
//
// Replaces, on the GPU, the numeric half of CHOLMOD:
//   * assembly of M = A_C W_C A_C' + diag(max(d_full - d_core, 0)) + delta I
//     (psipm_stable_preconditioner.py:373-395, psipm_active_pool.py:999-1006)
//     straight into the supernodal panels through the host-built scatter map,
//     with the unit-diagonal Jacobi scaling D M D (psipm_active_pool.py:850-853)
//     and the diagonal shift eta of the shift ladder (:855-885) fused in;
//   * the numeric factorisation (cholmod factorize / refactor, :668-678) as a
//     level-scheduled right-looking supernodal algorithm: per tree level one
//     launch factors every diagonal block in LDS (POTRF, plus the explicit
//     inverse of the block for the solves), one launch solves the panels
//     below (TRSM, one row per lane), one launch forms the 64x64 tiles of
//     L_b L_b' and scatters them into the ancestor panels through relative
//     indices staged in LDS (plain read-modify-write when a level holds one
//     supernode, hardware float atomics otherwise);
//   * the apply z = D P' L^-T L^-1 P D r (:903-904): forward solve one launch
//     per level (each block forms y_s = L_ss^-1 b_s from the stored inverse
//     and pushes L_b y_s into the ancestors), backward solve two launches per
//     level (coalesced column dot products L_b' x, then a fixed-order sum and
//     y_s -> L_ss^-T (.)), so the backward pass is deterministic.
// Panels, block inverses and solve vectors are FP64 (PSG_FP64) or FP32
// (PSG_MIXED); the assembly sums and the scaling are always FP64.
#include <algorithm>
#include <climits>
#include <cmath>

#include "ctx.hpp"

namespace psg {

constexpr int kW = 64;      // max supernode width handled by the panel kernels
constexpr int kRB = 256;    // rows per TRSM / solve unit
constexpr int kTile = 64;   // update tile edge

struct LevelRange {
  int node_begin, node_end;   // into level_nodes
  int unit_begin, unit_end;   // TRSM units (below rows only)
  int sunit_begin, sunit_end; // solve units (>= 1 per supernode)
  int tile_begin, tile_end;   // update tiles
};

struct FactorState {
  uint64_t serial = 0;
  int m = 0, nsuper = 0, nlev = 0;
  int precision = PSG_FP64;
  std::vector<LevelRange> levels;
  DevBuf<int> perm, sfirst, rows, snode_of, level_nodes, sparent;
  DevBuf<int64_t> rptr, poff, doff, diag_index, seg_target, seg_ptr, pair_a, pair_b;
  DevBuf<int> pair_col;
  DevBuf<int> unit_s, unit_k, sunit_s, sunit_k, sunit_first, tile_s, tile_i, tile_j;
  int64_t nsegs = 0;
  // numeric data (type depends on precision)
  DevBuf<double> panels64, dinv64, vb64, vy64, vx64, part64;
  DevBuf<float> panels32, dinv32, vb32, vy32, vx32, part32;
  DevBuf<double> dM;
  DevBuf<uint8_t> coremask;
  DevBuf<int> fail;
  int64_t npanel = 0, ndinv = 0;
  double nnz_L = 0.0;
  bool ready = false;
  std::vector<int> single_node_level;  // 1 if the level holds exactly one supernode
};

// ---------------------------------------------------------------------------
// assembly
// ---------------------------------------------------------------------------
__global__ void k_scaling(int m, const double* __restrict__ d_full, const double* __restrict__ d_core,
                          double delta, double* __restrict__ dM, double* __restrict__ dsc) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.x * blockDim.x) {
    const double dc = d_core[i];
    const double d = dc + fmax(d_full[i] - dc, 0.0) + delta;
    dM[i] = d;
    dsc[i] = 1.0 / sqrt(d > 0.0 ? d : 1.0);
  }
}

// panel[target] = dsc_r dsc_c sum_k w[col_k] a[pa_k] a[pb_k]  (fixed-order segment sum)
template <class T>
__global__ void k_assemble_offdiag(int64_t nseg, const int64_t* __restrict__ seg_target,
                                   const int64_t* __restrict__ seg_ptr, const int* __restrict__ pair_col,
                                   const int64_t* __restrict__ pair_a, const int64_t* __restrict__ pair_b,
                                   const double* __restrict__ aval, const int* __restrict__ arow,
                                   const double* __restrict__ w, const double* __restrict__ dsc,
                                   T* __restrict__ panels) {
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < nseg; t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t b = seg_ptr[t], e = seg_ptr[t + 1];
    double s = 0.0;
    for (int64_t k = b; k < e; ++k) s += w[pair_col[k]] * aval[pair_a[k]] * aval[pair_b[k]];
    s *= dsc[arow[pair_a[b]]] * dsc[arow[pair_b[b]]];
    panels[seg_target[t]] = (T)s;
  }
}

template <class T>
__global__ void k_assemble_diag(int m, const int* __restrict__ perm, const int64_t* __restrict__ diag_index,
                                const double* __restrict__ dM, const double* __restrict__ dsc, double eta,
                                T* __restrict__ panels) {
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < m; k += (int64_t)gridDim.x * blockDim.x) {
    const int i = perm[k];
    panels[diag_index[k]] = (T)(dM[i] * dsc[i] * dsc[i] + eta);
  }
}

// ---------------------------------------------------------------------------
// factorisation kernels
// ---------------------------------------------------------------------------
// POTRF of every diagonal block of a level in LDS; writes L_ss back to the
// panel and L_ss^-1 (column-major w x w) to dinv.  A non-positive pivot
// records its permuted column in *fail (atomic min).
template <class T>
__global__ void __launch_bounds__(kBlock)
k_potrf(const int* __restrict__ nodes, const int* __restrict__ sfirst, const int64_t* __restrict__ rptr,
        const int64_t* __restrict__ poff, const int64_t* __restrict__ doff, T* __restrict__ panels,
        T* __restrict__ dinv, int* __restrict__ fail) {
  __shared__ T sA[kW][kW + 1];
  __shared__ T tmp[kW];
  const int s = nodes[blockIdx.x];
  const int f = sfirst[s];
  const int w = sfirst[s + 1] - f;
  const int64_t nr = rptr[s + 1] - rptr[s];
  T* P = panels + poff[s];
  const int tid = threadIdx.x;
  for (int idx = tid; idx < w * w; idx += blockDim.x) {
    const int c = idx / w, r = idx % w;
    sA[r][c] = (r >= c) ? P[(int64_t)c * nr + r] : T(0);
  }
  __syncthreads();
  const int row = tid & (kW - 1);
  const int grp = tid / kW;  // 0..3
  bool bad = false;
  for (int j = 0; j < w; ++j) {
    __syncthreads();  // trailing update of step j-1 complete
    const T d = sA[j][j];
    const bool ok = d > T(0) && isfinite((double)d);
    if (!ok) bad = true;
    const T ljj = ok ? (T)sqrt((double)d) : T(NAN);
    __syncthreads();
    if (grp == 0 && row > j && row < w) sA[row][j] /= ljj;
    if (tid == 0) sA[j][j] = ljj;
    __syncthreads();
    if (row > j && row < w) {
      const T lij = sA[row][j];
      for (int k = j + 1 + grp; k <= row; k += kBlock / kW) sA[row][k] -= lij * sA[k][j];
    }
  }
  if (bad && tid == 0) atomicMin(fail, f);
  __syncthreads();
  for (int idx = tid; idx < w * w; idx += blockDim.x) {
    const int c = idx / w, r = idx % w;
    if (r >= c) P[(int64_t)c * nr + r] = sA[r][c];
  }
  __syncthreads();
  // in-place inverse of the lower-triangular block, last column first:
  // inv[j][j] = 1/l_jj, inv[j+1:, j] = -inv[j+1:, j+1:] l[j+1:, j] / l_jj
  for (int j = w - 1; j >= 0; --j) {
    if (tid > j && tid < w) {
      T acc = T(0);
      for (int k = j + 1; k <= tid; ++k) acc += sA[tid][k] * sA[k][j];
      tmp[tid] = acc;
    }
    if (tid == j) sA[j][j] = T(1) / sA[j][j];
    __syncthreads();
    if (tid > j && tid < w) sA[tid][j] = -tmp[tid] * sA[j][j];
    __syncthreads();
  }
  T* D = dinv + doff[s];
  for (int idx = tid; idx < w * w; idx += blockDim.x) {
    const int c = idx / w, r = idx % w;
    D[idx] = (r >= c) ? sA[r][c] : T(0);
  }
}

// Below-diagonal panel solve L_b = A_b L_ss^-T, one panel row per lane.
template <class T>
__global__ void __launch_bounds__(kBlock)
k_trsm(const int* __restrict__ unit_s, const int* __restrict__ unit_k, const int* __restrict__ sfirst,
       const int64_t* __restrict__ rptr, const int64_t* __restrict__ poff, T* __restrict__ panels) {
  __shared__ T sL[kW][kW + 1];
  const int s = unit_s[blockIdx.x];
  const int k = unit_k[blockIdx.x];
  const int w = sfirst[s + 1] - sfirst[s];
  const int64_t nr = rptr[s + 1] - rptr[s];
  T* P = panels + poff[s];
  for (int idx = threadIdx.x; idx < w * w; idx += blockDim.x) {
    const int c = idx / w, r = idx % w;
    sL[r][c] = (r >= c) ? P[(int64_t)c * nr + r] : T(0);
  }
  __syncthreads();
  const int64_t r = w + (int64_t)k * kRB + threadIdx.x;
  if (r >= nr) return;
  T x[kW];
#pragma unroll
  for (int c = 0; c < kW; ++c)
    if (c < w) x[c] = P[(int64_t)c * nr + r];
#pragma unroll
  for (int c = 0; c < kW; ++c) {
    if (c < w) {
      T acc = x[c];
#pragma unroll
      for (int j = 0; j < c; ++j) acc -= x[j] * sL[c][j];
      x[c] = acc / sL[c][c];
    }
  }
#pragma unroll
  for (int c = 0; c < kW; ++c)
    if (c < w) P[(int64_t)c * nr + r] = x[c];
}

// Update tile (I, J) of supernode s: C = L_I L_J' (64 x 64, k = w), scattered
// as panel[target] -= C into the ancestors that own the columns.
template <class T, bool ATOMIC>
__global__ void __launch_bounds__(kBlock)
k_update(const int* __restrict__ tile_s, const int* __restrict__ tile_i, const int* __restrict__ tile_j,
         const int* __restrict__ sfirst, const int64_t* __restrict__ rptr, const int* __restrict__ rows,
         const int64_t* __restrict__ poff, const int* __restrict__ snode_of, T* __restrict__ panels) {
  constexpr int kC = 32;  // k-chunk staged in LDS
  __shared__ T sI[kC][kTile + 1];
  __shared__ T sJ[kC][kTile + 1];
  __shared__ int rowI[kTile], colJ[kTile], tgt[kTile];
  __shared__ int64_t cbase[kTile];
  __shared__ int relpos[kTile];
  __shared__ int seg_start[kTile + 1];
  __shared__ int nseg_sh;
  const int s = tile_s[blockIdx.x];
  const int I = tile_i[blockIdx.x], J = tile_j[blockIdx.x];
  const int w = sfirst[s + 1] - sfirst[s];
  const int64_t nr = rptr[s + 1] - rptr[s];
  const int64_t nb = nr - w;
  const int* Rb = rows + rptr[s] + w;
  const T* P = panels + poff[s] + w;  // below part: element (r, c) at P[c*nr + r]
  const int nI = (int)min<int64_t>(kTile, nb - (int64_t)I * kTile);
  const int nJ = (int)min<int64_t>(kTile, nb - (int64_t)J * kTile);
  const int tid = threadIdx.x;
  if (tid < kTile) {
    rowI[tid] = (tid < nI) ? Rb[(int64_t)I * kTile + tid] : INT_MAX;
    if (tid < nJ) {
      const int col = Rb[(int64_t)J * kTile + tid];
      const int t = snode_of[col];
      colJ[tid] = col;
      tgt[tid] = t;
      cbase[tid] = poff[t] + (int64_t)(col - sfirst[t]) * (rptr[t + 1] - rptr[t]);
    } else {
      colJ[tid] = INT_MAX;
      tgt[tid] = -1;
      cbase[tid] = 0;
    }
  }
  __syncthreads();
  if (tid == 0) {
    int ns = 0;
    for (int b = 0; b < nJ; ++b)
      if (b == 0 || tgt[b] != tgt[b - 1]) seg_start[ns++] = b;
    seg_start[ns] = nJ;
    nseg_sh = ns;
  }
  // C tile: thread (tx, ty) owns rows ty*4.., cols tx*4..
  const int tx = tid % 16, ty = tid / 16;
  T acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = T(0);
  for (int c0 = 0; c0 < w; c0 += kC) {
    const int cw = min(kC, w - c0);
    __syncthreads();
    for (int idx = tid; idx < kC * kTile; idx += blockDim.x) {
      const int c = idx / kTile, a = idx % kTile;
      const bool inc = c < cw;
      sI[c][a] = (inc && a < nI) ? P[(int64_t)(c0 + c) * nr + (int64_t)I * kTile + a] : T(0);
      sJ[c][a] = (inc && a < nJ) ? P[(int64_t)(c0 + c) * nr + (int64_t)J * kTile + a] : T(0);
    }
    __syncthreads();
    for (int c = 0; c < cw; ++c) {
      T a[4], b[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) a[i] = sI[c][ty * 4 + i];
#pragma unroll
      for (int j = 0; j < 4; ++j) b[j] = sJ[c][tx * 4 + j];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] += a[i] * b[j];
    }
  }
  __syncthreads();
  const int nseg = nseg_sh;
  for (int sg = 0; sg < nseg; ++sg) {
    const int b0 = seg_start[sg], b1 = seg_start[sg + 1];
    const int t = tgt[b0];
    // relative positions of the I-tile rows inside the row structure of t
    if (tid < kTile) {
      int pos = -1;
      const int rI = rowI[tid];
      if (tid < nI && rI >= colJ[b0]) {
        const int* Rt = rows + rptr[t];
        int lo = 0, hi = (int)(rptr[t + 1] - rptr[t]);
        const int nrt = hi;
        while (lo < hi) {
          const int mid = (lo + hi) >> 1;
          if (Rt[mid] < rI) lo = mid + 1; else hi = mid;
        }
        pos = (lo < nrt && Rt[lo] == rI) ? lo : -1;  // -1: not in the structure (never for a valid plan)
      }
      relpos[tid] = pos;
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int b = tx * 4 + j;
      if (b < b0 || b >= b1) continue;
      const int col = colJ[b];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int a = ty * 4 + i;
        if (a >= nI || rowI[a] < col) continue;
        const int pos = relpos[a];
        if (pos < 0) continue;
        T* dst = panels + cbase[b] + pos;
        if (ATOMIC) unsafeAtomicAdd(dst, -acc[i][j]);
        else *dst -= acc[i][j];
      }
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// triangular solves
// ---------------------------------------------------------------------------
template <class T>
__global__ void k_perm_in(int m, const int* __restrict__ perm, const double* __restrict__ dsc,
                          const double* __restrict__ r, T* __restrict__ b) {
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < m; k += (int64_t)gridDim.x * blockDim.x) {
    const int i = perm[k];
    b[k] = (T)(dsc[i] * r[i]);
  }
}

template <class T>
__global__ void k_perm_out(int m, const int* __restrict__ perm, const double* __restrict__ dsc,
                           const T* __restrict__ x, double* __restrict__ z) {
  for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < m; k += (int64_t)gridDim.x * blockDim.x) {
    const int i = perm[k];
    z[i] = dsc[i] * (double)x[k];
  }
}

// forward unit (s, k): y_s = L_ss^-1 b_s (k == 0 stores it), then
// b[rows of block k] -= L_block y_s
template <class T, bool ATOMIC>
__global__ void __launch_bounds__(kBlock)
k_fwd(const int* __restrict__ sunit_s, const int* __restrict__ sunit_k, const int* __restrict__ sfirst,
      const int64_t* __restrict__ rptr, const int* __restrict__ rows, const int64_t* __restrict__ poff,
      const int64_t* __restrict__ doff, const T* __restrict__ panels, const T* __restrict__ dinv,
      T* __restrict__ b, T* __restrict__ y) {
  __shared__ T bs[kW], ys[kW];
  const int s = sunit_s[blockIdx.x];
  const int k = sunit_k[blockIdx.x];
  const int f = sfirst[s];
  const int w = sfirst[s + 1] - f;
  const int64_t nr = rptr[s + 1] - rptr[s];
  const int tid = threadIdx.x;
  if (tid < w) bs[tid] = b[f + tid];
  __syncthreads();
  if (tid < w) {
    const T* D = dinv + doff[s];
    T acc = T(0);
    for (int c = 0; c <= tid; ++c) acc += D[(int64_t)c * w + tid] * bs[c];
    ys[tid] = acc;
    if (k == 0) y[f + tid] = acc;
  }
  __syncthreads();
  const int64_t r = w + (int64_t)k * kRB + tid;
  if (r >= nr) return;
  const T* P = panels + poff[s];
  T acc = T(0);
  for (int c = 0; c < w; ++c) acc += P[(int64_t)c * nr + r] * ys[c];
  const int row = rows[rptr[s] + r];
  if (ATOMIC) unsafeAtomicAdd(b + row, -acc);
  else b[row] -= acc;
}

// backward partials of unit u = (s, k): part[u][c] = sum_{rows of block} L[r][c] x[row]
template <class T>
__global__ void __launch_bounds__(kBlock)
k_bwd_partial(int ubase, const int* __restrict__ sunit_s, const int* __restrict__ sunit_k,
              const int* __restrict__ sfirst, const int64_t* __restrict__ rptr, const int* __restrict__ rows,
              const int64_t* __restrict__ poff, const T* __restrict__ panels, const T* __restrict__ x,
              T* __restrict__ part) {
  __shared__ T xr[kRB];
  const int u = ubase + blockIdx.x;
  const int s = sunit_s[u];
  const int k = sunit_k[u];
  const int w = sfirst[s + 1] - sfirst[s];
  const int64_t nr = rptr[s + 1] - rptr[s];
  const int64_t r0 = w + (int64_t)k * kRB;
  const int nrow = (int)max<int64_t>(0, min<int64_t>(kRB, nr - r0));
  const int tid = threadIdx.x;
  if (tid < nrow) xr[tid] = x[rows[rptr[s] + r0 + tid]];
  __syncthreads();
  const T* P = panels + poff[s];
  const int lane = tid & (kWave - 1), wv = tid / kWave;
  for (int c = wv; c < w; c += kBlock / kWave) {
    double acc = 0.0;
    for (int i = lane; i < nrow; i += kWave) acc += (double)(P[(int64_t)c * nr + r0 + i] * xr[i]);
    acc = wave_sum(acc);
    if (lane == 0) part[(int64_t)u * kW + c] = (T)acc;
  }
}

// backward finish of supernode s: x_s = L_ss^-T (y_s - sum_k part[u_k])
template <class T>
__global__ void __launch_bounds__(kBlock)
k_bwd_finish(const int* __restrict__ nodes, const int* __restrict__ sunit_range, const int* __restrict__ sfirst,
             const int64_t* __restrict__ doff, const T* __restrict__ dinv, const T* __restrict__ part,
             const T* __restrict__ y, T* __restrict__ x) {
  __shared__ T cs[kW];
  const int s = nodes[blockIdx.x];
  const int f = sfirst[s];
  const int w = sfirst[s + 1] - f;
  const int u0 = sunit_range[2 * s], u1 = sunit_range[2 * s + 1];
  const int tid = threadIdx.x;
  if (tid < w) {
    T acc = y[f + tid];
    for (int u = u0; u < u1; ++u) acc -= part[(int64_t)u * kW + tid];
    cs[tid] = acc;
  }
  __syncthreads();
  if (tid < w) {
    const T* D = dinv + doff[s];
    T acc = T(0);
    for (int j = tid; j < w; ++j) acc += D[(int64_t)tid * w + j] * cs[j];
    x[f + tid] = acc;
  }
}

// ---------------------------------------------------------------------------
// host orchestration
// ---------------------------------------------------------------------------
static void upload_plan(psg_ctx* c, const Plan& P) {
  FactorState* fs = c->fs;
  hipStream_t s = c->stream;
  fs->serial = P.serial;
  fs->m = P.m;
  fs->nsuper = P.nsuper;
  fs->nlev = P.nlevels;
  fs->precision = c->precision;
  if (P.max_width > kW) throw Error(PSG_ERR_ARGUMENT, "plan supernode width exceeds 64");
  std::vector<int64_t> doff(P.nsuper + 1, 0);
  for (int k = 0; k < P.nsuper; ++k) {
    const int64_t w = P.sfirst[k + 1] - P.sfirst[k];
    doff[k + 1] = doff[k] + w * w;
  }
  // unit and tile lists per level
  std::vector<int> unit_s, unit_k, sunit_s, sunit_k, tile_s, tile_i, tile_j;
  std::vector<int> sunit_first(P.nsuper + 1, 0);
  fs->levels.assign(P.nlevels, {});
  fs->single_node_level.assign(P.nlevels, 0);
  for (int l = 0; l < P.nlevels; ++l) {
    LevelRange& L = fs->levels[l];
    L.node_begin = P.level_ptr[l];
    L.node_end = P.level_ptr[l + 1];
    fs->single_node_level[l] = (L.node_end - L.node_begin) == 1;
    L.unit_begin = (int)unit_s.size();
    L.sunit_begin = (int)sunit_s.size();
    L.tile_begin = (int)tile_s.size();
    for (int q = L.node_begin; q < L.node_end; ++q) {
      const int sn = P.level_nodes[q];
      const int w = P.sfirst[sn + 1] - P.sfirst[sn];
      const int64_t nb = (P.rptr[sn + 1] - P.rptr[sn]) - w;
      const int64_t nblk = (nb + kRB - 1) / kRB;
      for (int64_t k = 0; k < nblk; ++k) {
        unit_s.push_back(sn);
        unit_k.push_back((int)k);
      }
      sunit_first[sn] = (int)sunit_s.size();
      for (int64_t k = 0; k < std::max<int64_t>(1, nblk); ++k) {
        sunit_s.push_back(sn);
        sunit_k.push_back((int)k);
      }
      const int64_t nt = (nb + kTile - 1) / kTile;
      for (int64_t i = 0; i < nt; ++i)
        for (int64_t j = 0; j <= i; ++j) {
          tile_s.push_back(sn);
          tile_i.push_back((int)i);
          tile_j.push_back((int)j);
        }
    }
    L.unit_end = (int)unit_s.size();
    L.sunit_end = (int)sunit_s.size();
    L.tile_end = (int)tile_s.size();
  }
  // [begin, end) of the solve units of each supernode
  std::vector<int> sunit_range(2 * P.nsuper);
  {
    std::vector<int> count(P.nsuper, 0);
    for (int sn : sunit_s) count[sn]++;
    for (int sn = 0; sn < P.nsuper; ++sn) {
      sunit_range[2 * sn] = sunit_first[sn];
      sunit_range[2 * sn + 1] = sunit_first[sn] + count[sn];
    }
  }
  fs->perm.upload(P.perm, s);
  fs->sfirst.upload(P.sfirst, s);
  fs->rows.upload(P.rows, s);
  fs->snode_of.upload(P.snode_of, s);
  fs->level_nodes.upload(P.level_nodes, s);
  fs->rptr.upload(P.rptr, s);
  fs->poff.upload(P.poff, s);
  fs->doff.upload(doff, s);
  fs->diag_index.upload(P.diag_index, s);
  fs->seg_target.upload(P.seg_target, s);
  fs->seg_ptr.upload(P.seg_ptr, s);
  fs->pair_a.upload(P.pair_a, s);
  fs->pair_b.upload(P.pair_b, s);
  fs->pair_col.upload(P.pair_col, s);
  fs->unit_s.upload(unit_s, s);
  fs->unit_k.upload(unit_k, s);
  fs->sunit_s.upload(sunit_s, s);
  fs->sunit_k.upload(sunit_k, s);
  fs->sunit_first.upload(sunit_range, s);
  fs->tile_s.upload(tile_s, s);
  fs->tile_i.upload(tile_i, s);
  fs->tile_j.upload(tile_j, s);
  fs->nsegs = (int64_t)P.seg_target.size();
  fs->npanel = P.nnz_panels;
  fs->ndinv = doff[P.nsuper];
  fs->nnz_L = (double)P.nnz_L;
  const size_t npart = (size_t)std::max<size_t>(1, sunit_s.size()) * kW;
  if (c->precision == PSG_MIXED) {
    fs->panels32.alloc(std::max<int64_t>(1, fs->npanel));
    fs->dinv32.alloc(std::max<int64_t>(1, fs->ndinv));
    for (auto* b : {&fs->vb32, &fs->vy32, &fs->vx32}) b->alloc(P.m);
    fs->part32.alloc(npart);
    fs->panels64.release();
    fs->dinv64.release();
  } else {
    fs->panels64.alloc(std::max<int64_t>(1, fs->npanel));
    fs->dinv64.alloc(std::max<int64_t>(1, fs->ndinv));
    for (auto* b : {&fs->vb64, &fs->vy64, &fs->vx64}) b->alloc(P.m);
    fs->part64.alloc(npart);
    fs->panels32.release();
    fs->dinv32.release();
  }
  fs->dM.alloc(P.m);
  fs->fail.alloc(1);
  std::vector<uint8_t> mask(c->n, 0);
  for (int j : P.core) {
    if (j < 0 || j >= c->n) throw Error(PSG_ERR_ARGUMENT, "plan core column outside the context's A");
    mask[j] = 1;
  }
  fs->coremask.upload(mask, s);
  HIP_CHECK(hipStreamSynchronize(s));
}

template <class T>
static int factor_run(psg_ctx* c, const Plan& P, double eta, T* panels, T* dinv) {
  FactorState* fs = c->fs;
  hipStream_t s = c->stream;
  // exact dropped diagonal: d_full over all columns, d_core over the active ones
  const int m = c->m;
  diag_terms_device(c, fs->coremask.p, c->d_full.p, c->d_core.p);
  hipLaunchKernelGGL(k_scaling, dim3(grid_for(m, kBlock, 4096)), dim3(kBlock), 0, s, m, c->d_full.p,
                     c->d_core.p, c->delta, fs->dM.p, c->dsc.p);
  HIP_LAUNCH_CHECK();
  HIP_CHECK(hipMemsetAsync(panels, 0, (size_t)fs->npanel * sizeof(T), s));
  if (fs->nsegs)
    hipLaunchKernelGGL((k_assemble_offdiag<T>), dim3(grid_for(fs->nsegs, kBlock, 16384)), dim3(kBlock), 0, s,
                       fs->nsegs, fs->seg_target.p, fs->seg_ptr.p, fs->pair_col.p, fs->pair_a.p, fs->pair_b.p,
                       c->cval.p, c->crow.p, c->w.p, c->dsc.p, panels);
  HIP_LAUNCH_CHECK();
  hipLaunchKernelGGL((k_assemble_diag<T>), dim3(grid_for(m, kBlock, 4096)), dim3(kBlock), 0, s, m, fs->perm.p,
                     fs->diag_index.p, fs->dM.p, c->dsc.p, eta, panels);
  HIP_LAUNCH_CHECK();
  const int big = INT_MAX;
  HIP_CHECK(hipMemcpyAsync(fs->fail.p, &big, sizeof(int), hipMemcpyHostToDevice, s));
  for (int l = 0; l < fs->nlev; ++l) {
    const LevelRange& L = fs->levels[l];
    const int nn = L.node_end - L.node_begin;
    hipLaunchKernelGGL((k_potrf<T>), dim3(nn), dim3(kBlock), 0, s, fs->level_nodes.p + L.node_begin,
                       fs->sfirst.p, fs->rptr.p, fs->poff.p, fs->doff.p, panels, dinv, fs->fail.p);
    HIP_LAUNCH_CHECK();
    const int nu = L.unit_end - L.unit_begin;
    if (nu) {
      hipLaunchKernelGGL((k_trsm<T>), dim3(nu), dim3(kBlock), 0, s, fs->unit_s.p + L.unit_begin,
                         fs->unit_k.p + L.unit_begin, fs->sfirst.p, fs->rptr.p, fs->poff.p, panels);
      HIP_LAUNCH_CHECK();
    }
    const int nt = L.tile_end - L.tile_begin;
    if (nt) {
      if (fs->single_node_level[l])
        hipLaunchKernelGGL((k_update<T, false>), dim3(nt), dim3(kBlock), 0, s, fs->tile_s.p + L.tile_begin,
                           fs->tile_i.p + L.tile_begin, fs->tile_j.p + L.tile_begin, fs->sfirst.p, fs->rptr.p,
                           fs->rows.p, fs->poff.p, fs->snode_of.p, panels);
      else
        hipLaunchKernelGGL((k_update<T, true>), dim3(nt), dim3(kBlock), 0, s, fs->tile_s.p + L.tile_begin,
                           fs->tile_i.p + L.tile_begin, fs->tile_j.p + L.tile_begin, fs->sfirst.p, fs->rptr.p,
                           fs->rows.p, fs->poff.p, fs->snode_of.p, panels);
      HIP_LAUNCH_CHECK();
    }
  }
  int failcol = INT_MAX;
  HIP_CHECK(hipMemcpyAsync(&failcol, fs->fail.p, sizeof(int), hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipStreamSynchronize(s));
  return failcol == INT_MAX ? 0 : failcol + 1;
}

template <class T>
static void apply_run(psg_ctx* c, const double* r, double* z, const T* panels, const T* dinv, T* vb, T* vy,
                      T* vx, T* part) {
  FactorState* fs = c->fs;
  hipStream_t s = c->stream;
  const int m = c->m;
  hipLaunchKernelGGL((k_perm_in<T>), dim3(grid_for(m, kBlock, 4096)), dim3(kBlock), 0, s, m, fs->perm.p, c->dsc.p,
                     r, vb);
  HIP_LAUNCH_CHECK();
  for (int l = 0; l < fs->nlev; ++l) {
    const LevelRange& L = fs->levels[l];
    const int nu = L.sunit_end - L.sunit_begin;
    if (fs->single_node_level[l])
      hipLaunchKernelGGL((k_fwd<T, false>), dim3(nu), dim3(kBlock), 0, s, fs->sunit_s.p + L.sunit_begin,
                         fs->sunit_k.p + L.sunit_begin, fs->sfirst.p, fs->rptr.p, fs->rows.p, fs->poff.p,
                         fs->doff.p, panels, dinv, vb, vy);
    else
      hipLaunchKernelGGL((k_fwd<T, true>), dim3(nu), dim3(kBlock), 0, s, fs->sunit_s.p + L.sunit_begin,
                         fs->sunit_k.p + L.sunit_begin, fs->sfirst.p, fs->rptr.p, fs->rows.p, fs->poff.p,
                         fs->doff.p, panels, dinv, vb, vy);
    HIP_LAUNCH_CHECK();
  }
  for (int l = fs->nlev - 1; l >= 0; --l) {
    const LevelRange& L = fs->levels[l];
    const int nu = L.sunit_end - L.sunit_begin;
    hipLaunchKernelGGL((k_bwd_partial<T>), dim3(nu), dim3(kBlock), 0, s, L.sunit_begin, fs->sunit_s.p,
                       fs->sunit_k.p, fs->sfirst.p, fs->rptr.p, fs->rows.p, fs->poff.p, panels, vx, part);
    HIP_LAUNCH_CHECK();
    const int nn = L.node_end - L.node_begin;
    hipLaunchKernelGGL((k_bwd_finish<T>), dim3(nn), dim3(kBlock), 0, s, fs->level_nodes.p + L.node_begin,
                       fs->sunit_first.p, fs->sfirst.p, fs->doff.p, dinv, part, vy, vx);
    HIP_LAUNCH_CHECK();
  }
  hipLaunchKernelGGL((k_perm_out<T>), dim3(grid_for(m, kBlock, 4096)), dim3(kBlock), 0, s, m, fs->perm.p,
                     c->dsc.p, vx, z);
  HIP_LAUNCH_CHECK();
}

void factor_destroy(psg_ctx* c) {
  delete c->fs;
  c->fs = nullptr;
}

bool factor_ready(const psg_ctx* c) { return c->fs && c->fs->ready; }

int factor_numeric(psg_ctx* c, const Plan& P, double eta, double* nnz_factor) {
  if (!c->fs) c->fs = new FactorState;
  if (P.m != c->m) throw Error(PSG_ERR_ARGUMENT, "plan dimension does not match the context");
  if (c->fs->serial != P.serial || c->fs->precision != c->precision) upload_plan(c, P);
  c->fs->ready = false;
  HIP_CHECK(hipEventRecord(c->ev0, c->stream));
  int status;
  if (c->precision == PSG_MIXED)
    status = factor_run<float>(c, P, eta, c->fs->panels32.p, c->fs->dinv32.p);
  else
    status = factor_run<double>(c, P, eta, c->fs->panels64.p, c->fs->dinv64.p);
  HIP_CHECK(hipEventRecord(c->ev1, c->stream));
  HIP_CHECK(hipEventSynchronize(c->ev1));
  float ms = 0.f;
  HIP_CHECK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
  c->last_factor_ms = ms;
  if (nnz_factor) *nnz_factor = c->fs->nnz_L;
  c->fs->ready = (status == 0);
  return status;
}

void factor_apply(psg_ctx* c, const double* r, double* z) {
  FactorState* fs = c->fs;
  if (c->precision == PSG_MIXED)
    apply_run<float>(c, r, z, fs->panels32.p, fs->dinv32.p, fs->vb32.p, fs->vy32.p, fs->vx32.p, fs->part32.p);
  else
    apply_run<double>(c, r, z, fs->panels64.p, fs->dinv64.p, fs->vb64.p, fs->vy64.p, fs->vx64.p, fs->part64.p);
}

}  // namespace psg

using namespace psg;

extern "C" int psg_factor(psg_ctx* c, const psg_plan* plan, double eta, double* nnz_factor) {
  return guarded([&]() -> int {
    if (!c || !plan) throw Error(PSG_ERR_ARGUMENT, "psg_factor: null argument");
    if (!c->have_weights) throw Error(PSG_ERR_STATE, "psg_factor: weights not set");
    return factor_numeric(c, plan->plan, eta, nnz_factor);
  });
}
