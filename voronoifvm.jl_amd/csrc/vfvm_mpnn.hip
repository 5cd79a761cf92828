// Learned Newton initial guess: a message-passing network on the Voronoi edge
// graph of a system, with hand-written forward and backward kernels.  Dense
// layers run on the matrix cores (v_mfma_f32_16x16x16_bf16: bfloat16 operands,
// float32 accumulation); aggregation, activations, loss and Adam are plain
// kernels.  The network (row r = sample s, node i of the system's grid):
//
//   x_r   : 8 input features; channels 8..15 are multi-scale context: the features
//           a = x_1 (asinh(f/2)) and x_0 (f/5) smoothed by K = 16, 64, 160, 324 steps
//           of y <- (y + W y) / 2 on the graph (Gaussian-like, width sqrt(K)/2 edges)
//   h0    = relu(bf(x) We + be)                                   64 wide
//   l = 0..L-1:
//     a_l = sum_j w_ij h_l[j] - h_l[i]     w_ij = sigma_ij/h_ij normalised per row
//     c_l = [bf(h_l), bf(a_l)]                                    128 wide, bfloat16
//     r_l = relu(c_l W_l + b_l),  h_{l+1} = h_l + r_l
//   u     = h_L wd + bd
//
// Parameters live in one float32 vector (layout: We 16x64, be, [W_l 128x64, b_l]*L,
// wd 64, bd) with bfloat16 copies (W_l and its transpose) refreshed after every
// optimiser step.  Training minimises the mean squared error to converged Newton
// solutions.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <random>
#include <vector>

#include "vfvm_sys.h"

using namespace vfvm;

namespace {

constexpr int NFEAT = 8;   // host-side input features
constexpr int KPAD = 16;   // padded input width (one MFMA k-step)
constexpr int WID = 64;    // hidden width
constexpr int CW = 2 * WID;  // concatenated width
constexpr int TB = 256;    // threads per block (4 waves)

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef uint16_t bf16_t;

__device__ __forceinline__ bf16_t f2bf(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7fffffffu) > 0x7f800000u) return (bf16_t)((u >> 16) | 0x40u);  // quiet NaN
  u += 0x7fffu + ((u >> 16) & 1u);                                          // round to nearest even
  return (bf16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(bf16_t h) { return __uint_as_float(((uint32_t)h) << 16); }

// Workgroups are dealt to the 8 XCDs round-robin (workgroup b runs on XCD b % 8).  For a
// grid of G = 8k workgroups, logical block (b % 8) * k + b / 8 gives each XCD one
// contiguous range, so the graph neighbours a block gathers (nearby rows) are mostly
// in the same XCD's L2.
__device__ __forceinline__ int64_t xcd_block(int64_t b, int64_t G) { return (b & 7) * (G >> 3) + (b >> 3); }

// ---------------------------------------------------------------- GEMM -----
// C[M x ND] = A[M x KD] * B[KD x ND], A and B bfloat16 row-major, float32
// accumulation on the matrix cores.  Block = 4 waves = 64 rows; wave w owns
// rows 16w..16w+15 and all ND columns (ND/16 accumulator tiles).  Both
// operands are staged in LDS (A rows padded by 8 to spread banks).
// Operand layout of v_mfma_f32_16x16x16_bf16 for lane l (g = l / 16, c = l % 16):
//   A: A[c][4g + j], B: B[4g + j][c], D: D[4g + j][c], j = 0..3.
enum { EPI_ENC = 0, EPI_RES = 1, EPI_DCAT = 2 };

struct GemmEpi {
  const float *bias;  // ND (ENC, RES)
  float *hf;          // rows x 64 hidden state (ENC: written, RES: updated in place)
  float *rbuf;        // rows x 64 relu(z) kept for the backward pass (RES)
  bf16_t *cat;        // rows x 128: ENC/RES write bf(h) into cols 0..63 of the next layer input
  float *dh;          // DCAT: dh += dC[:, 0:64]
  float *da;          // DCAT: da  = dC[:, 64:128]
};

template <int KD, int ND, int EPI>
__global__ void __launch_bounds__(TB) k_gemm(int64_t M, const bf16_t *__restrict__ A, const bf16_t *__restrict__ B,
                                             GemmEpi e) {
  constexpr int ASTR = KD + 8;
  __shared__ bf16_t As[64 * ASTR];
  __shared__ bf16_t Bs[KD * ND];
  const int tid = threadIdx.x, w = tid >> 6, l = tid & 63, g = l >> 4, c = l & 15;
  const int64_t row0 = (int64_t)blockIdx.x * 64;
  for (int q = tid; q < KD * ND / 4; q += TB) {  // B: 4 bf16 per load
    const uint64_t v = reinterpret_cast<const uint64_t *>(B)[q];
    reinterpret_cast<uint64_t *>(Bs)[q] = v;
  }
  for (int q = tid; q < 64 * KD / 4; q += TB) {  // A tile rows, zero beyond M
    const int r = q / (KD / 4), kq = q % (KD / 4);
    uint64_t v = 0;
    if (row0 + r < M) v = reinterpret_cast<const uint64_t *>(A + (row0 + r) * KD)[kq];
    *reinterpret_cast<uint64_t *>(&As[r * ASTR + kq * 4]) = v;
  }
  __syncthreads();
  f32x4 acc[ND / 16];
#pragma unroll
  for (int t = 0; t < ND / 16; t++) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int k0 = 0; k0 < KD; k0 += 16) {
    bf16x4 a;
    const bf16_t *ap = &As[(w * 16 + c) * ASTR + k0 + 4 * g];
#pragma unroll
    for (int j = 0; j < 4; j++) a[j] = (short)ap[j];
#pragma unroll
    for (int t = 0; t < ND / 16; t++) {
      bf16x4 b;
#pragma unroll
      for (int j = 0; j < 4; j++) b[j] = (short)Bs[(k0 + 4 * g + j) * ND + t * 16 + c];
      acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[t], 0, 0, 0);
    }
  }
#pragma unroll
  for (int t = 0; t < ND / 16; t++) {
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const int64_t row = row0 + w * 16 + 4 * g + j;
      const int col = t * 16 + c;
      if (row >= M) continue;
      const float v = acc[t][j];
      if constexpr (EPI == EPI_ENC) {
        const float h = fmaxf(v + e.bias[col], 0.f);
        e.hf[row * WID + col] = h;
        if (e.cat) e.cat[row * CW + col] = f2bf(h);
      } else if constexpr (EPI == EPI_RES) {
        const float r = fmaxf(v + e.bias[col], 0.f);
        e.rbuf[row * WID + col] = r;
        const float hn = e.hf[row * WID + col] + r;
        e.hf[row * WID + col] = hn;
        if (e.cat) e.cat[row * CW + col] = f2bf(hn);
      } else {
        if (col < WID)
          e.dh[row * WID + col] += v;
        else
          e.da[row * WID + col - WID] = v;
      }
    }
  }
}

// Weight gradient: P_b[KD x 64] = sum over the block's rows of A[rows x KD]^T Z[rows x 64],
// A and Z bfloat16.  Reduction over rows on the matrix cores (16 rows per step):
// A-operand A'[k][m] = A[m][k] (m = row), B-operand Z[m][n].  Output tiles
// (KD/16) x 4 are split over the 4 waves.
template <int KD>
__global__ void __launch_bounds__(TB) k_wgrad(int64_t M, int64_t rows_per_block, const bf16_t *__restrict__ A,
                                              const bf16_t *__restrict__ Z, float *__restrict__ partials) {
  constexpr int NT = (KD / 16) * (WID / 16);
  constexpr int PER = (NT + 3) / 4;
  constexpr int ASTR = KD + 8, ZSTR = WID + 8;
  __shared__ bf16_t As[64 * ASTR];
  __shared__ bf16_t Zs[64 * ZSTR];
  const int tid = threadIdx.x, w = tid >> 6, l = tid & 63, g = l >> 4, c = l & 15;
  f32x4 acc[PER];
#pragma unroll
  for (int p = 0; p < PER; p++) acc[p] = f32x4{0.f, 0.f, 0.f, 0.f};
  const int64_t r_begin = (int64_t)blockIdx.x * rows_per_block;
  const int64_t r_end = std::min<int64_t>(M, r_begin + rows_per_block);
  for (int64_t r0 = r_begin; r0 < r_end; r0 += 64) {
    __syncthreads();
    for (int q = tid; q < 64 * KD / 4; q += TB) {
      const int r = q / (KD / 4), kq = q % (KD / 4);
      uint64_t v = 0;
      if (r0 + r < r_end) v = reinterpret_cast<const uint64_t *>(A + (r0 + r) * KD)[kq];
      *reinterpret_cast<uint64_t *>(&As[r * ASTR + kq * 4]) = v;
    }
    for (int q = tid; q < 64 * WID / 4; q += TB) {
      const int r = q / (WID / 4), kq = q % (WID / 4);
      uint64_t v = 0;
      if (r0 + r < r_end) v = reinterpret_cast<const uint64_t *>(Z + (r0 + r) * WID)[kq];
      *reinterpret_cast<uint64_t *>(&Zs[r * ZSTR + kq * 4]) = v;
    }
    __syncthreads();
#pragma unroll
    for (int m0 = 0; m0 < 64; m0 += 16) {
#pragma unroll
      for (int p = 0; p < PER; p++) {
        const int tile = w * PER + p;
        if (tile >= NT) break;
        const int kt = tile / (WID / 16), nt = tile % (WID / 16);
        bf16x4 a, b;
#pragma unroll
        for (int j = 0; j < 4; j++) {
          a[j] = (short)As[(m0 + 4 * g + j) * ASTR + kt * 16 + c];
          b[j] = (short)Zs[(m0 + 4 * g + j) * ZSTR + nt * 16 + c];
        }
        acc[p] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[p], 0, 0, 0);
      }
    }
  }
  float *out = partials + (size_t)blockIdx.x * KD * WID;
#pragma unroll
  for (int p = 0; p < PER; p++) {
    const int tile = w * PER + p;
    if (tile >= NT) break;
    const int kt = tile / (WID / 16), nt = tile % (WID / 16);
#pragma unroll
    for (int j = 0; j < 4; j++) out[(kt * 16 + 4 * g + j) * WID + nt * 16 + c] = acc[p][j];
  }
}

// grad[k] = sum over blocks of partials[b][k]
__global__ void k_sum_partials(int nblocks, int len, const float *__restrict__ partials, float *__restrict__ grad) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= len) return;
  float acc = 0.f;
  for (int b = 0; b < nblocks; b++) acc += partials[(size_t)b * len + k];
  grad[k] = acc;
}

// ------------------------------------------------------------ graph ops ----
// a_i = sum_j w_ij h_j - h_i for row r = s*n + i; one wave (64 features) per row
__global__ void __launch_bounds__(TB) k_aggregate(int64_t rows, int64_t n, const int32_t *__restrict__ gptr,
                                                  const int32_t *__restrict__ gcol, const float *__restrict__ gw,
                                                  const float *__restrict__ hf, bf16_t *__restrict__ cat) {
  const int64_t row = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int f = threadIdx.x & 63;
  if (row >= rows) return;
  const int64_t s = row / n, i = row % n;
  float acc = 0.f;
  for (int32_t k = gptr[i]; k < gptr[i + 1]; k++) acc += gw[k] * hf[(s * n + gcol[k]) * WID + f];
  acc -= hf[row * WID + f];
  cat[row * CW + WID + f] = f2bf(acc);
}

// dh_k += sum_{i in N(k)} w_ik da_i - da_k  (transpose weights gwt on k's entries)
__global__ void __launch_bounds__(TB) k_aggregate_t(int64_t rows, int64_t n, const int32_t *__restrict__ gptr,
                                                    const int32_t *__restrict__ gcol, const float *__restrict__ gwt,
                                                    const float *__restrict__ da, float *__restrict__ dh) {
  const int64_t row = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int f = threadIdx.x & 63;
  if (row >= rows) return;
  const int64_t s = row / n, k = row % n;
  float acc = 0.f;
  for (int32_t q = gptr[k]; q < gptr[k + 1]; q++) acc += gwt[q] * da[(s * n + gcol[q]) * WID + f];
  dh[row * WID + f] += acc - da[row * WID + f];
}

// One hidden layer for inference, aggregation fused into the GEMM: for a tile of 64
// consecutive rows, c = [bf(h_l), bf(sum_j w_ij h_l[j] - h_l[i])] is built in LDS
// (same arithmetic as k_aggregate), multiplied on the matrix cores by W_l, and
// h_{l+1} = h_l + relu(c W_l + b_l) is written to `hout` (a buffer other than `hin`:
// neighbouring tiles still read h_l).  With `wd` set (last layer) the tile writes
// u = h_{l+1} . wd + bd instead.  Blocks loop over tiles; W_l sits in LDS in operand
// order (one 8-byte fragment per lane, k-step and column tile).  Each thread stages
// four (row, 4-column chunk) items and loads the first NB0 neighbours of two items
// before summing them, so 8 neighbour rows are in flight per thread.
constexpr int NB0 = 4;
__global__ void __launch_bounds__(TB, 4) k_layer_infer(int64_t rows, int64_t n, const int32_t *__restrict__ gptr,
                                                    const int32_t *__restrict__ gcol, const float *__restrict__ gw,
                                                    const float *__restrict__ hin, const bf16_t *__restrict__ W,
                                                    const float *__restrict__ bias, float *__restrict__ hout,
                                                    const float *__restrict__ wd, const float *__restrict__ bd,
                                                    float *__restrict__ u) {
  constexpr int ASTR = CW + 8;  // bf16 per LDS row of the layer input
  constexpr int KS = CW / 16, NTL = WID / 16;
  __shared__ bf16_t As[64 * ASTR];
  __shared__ uint64_t Bf[KS * NTL * 64];
  const int tid = threadIdx.x, w = tid >> 6, l = tid & 63, g = l >> 4, c = l & 15;
  for (int q = tid; q < KS * NTL * 64; q += TB) {
    const int k = q / (NTL * 64), t = (q / 64) % NTL, ln = q % 64, gg = ln >> 4, cc = ln & 15;
    uint64_t v = 0;
#pragma unroll
    for (int jj = 0; jj < 4; jj++) v |= (uint64_t)W[(k * 16 + 4 * gg + jj) * WID + t * 16 + cc] << (16 * jj);
    Bf[q] = v;
  }
  float bcol[NTL], wdc[NTL];
#pragma unroll
  for (int t = 0; t < NTL; t++) {
    bcol[t] = bias[t * 16 + c];
    wdc[t] = wd ? wd[t * 16 + c] : 0.f;
  }
  // XCD x takes the contiguous tile range [x T, (x + 1) T), T = ceil(ntiles / 8), in
  // steps of the grid's blocks per XCD
  const int64_t ntiles = (rows + 63) / 64, per = (ntiles + 7) / 8, G8 = gridDim.x / 8;
  const int64_t t_end = std::min<int64_t>(ntiles, (int64_t)(blockIdx.x & 7) * per + per);
  for (int64_t tile = (int64_t)(blockIdx.x & 7) * per + (blockIdx.x >> 3); tile < t_end; tile += G8) {
    const int64_t row0 = tile * 64;
    __syncthreads();  // Bf is written; the previous tile's As reads are done
    const float4 z4 = make_float4(0.f, 0.f, 0.f, 0.f);
    float4 hv[4], av[4];
    int64_t base[4];
    int32_t kb[4], ke[4];
#pragma unroll
    for (int it = 0; it < 4; it++) {
      const int q = tid + it * TB, r = q >> 4, ch = q & 15;
      const int64_t row = row0 + r;
      hv[it] = z4;
      av[it] = z4;
      base[it] = ch * 4;
      kb[it] = 0;
      ke[it] = 0;
      if (row < rows) {
        const int64_t smp = row / n, i = row - smp * n;
        base[it] = smp * n * WID + ch * 4;
        hv[it] = *reinterpret_cast<const float4 *>(hin + row * WID + ch * 4);
        kb[it] = gptr[i];
        ke[it] = gptr[i + 1];
      }
    }
#pragma unroll
    for (int p2 = 0; p2 < 4; p2 += 2) {
      // first NB0 neighbours of two items: indices clamped to entry 0 where absent
      // (a valid node of the same sample), skipped in the sum
      float4 v[2][NB0];
      float wk[2][NB0];
#pragma unroll
      for (int it = 0; it < 2; it++)
#pragma unroll
        for (int jj = 0; jj < NB0; jj++) {
          const int32_t kk = kb[p2 + it] + jj;
          const int32_t kc = kk < ke[p2 + it] ? kk : 0;
          wk[it][jj] = gw[kc];
          v[it][jj] = *reinterpret_cast<const float4 *>(hin + base[p2 + it] + (int64_t)gcol[kc] * WID);
        }
#pragma unroll
      for (int it = 0; it < 2; it++)
#pragma unroll
        for (int jj = 0; jj < NB0; jj++)
          if (kb[p2 + it] + jj < ke[p2 + it]) {
            av[p2 + it].x += wk[it][jj] * v[it][jj].x;
            av[p2 + it].y += wk[it][jj] * v[it][jj].y;
            av[p2 + it].z += wk[it][jj] * v[it][jj].z;
            av[p2 + it].w += wk[it][jj] * v[it][jj].w;
          }
    }
#pragma unroll
    for (int it = 0; it < 4; it++) {
      for (int32_t k = kb[it] + NB0; k < ke[it]; k++) {
        const float wv = gw[k];
        const float4 v = *reinterpret_cast<const float4 *>(hin + base[it] + (int64_t)gcol[k] * WID);
        av[it].x += wv * v.x;
        av[it].y += wv * v.y;
        av[it].z += wv * v.z;
        av[it].w += wv * v.w;
      }
      const int q = tid + it * TB, r = q >> 4, ch = q & 15;
      bf16_t *ar = &As[r * ASTR + ch * 4];
      ar[0] = f2bf(hv[it].x);
      ar[1] = f2bf(hv[it].y);
      ar[2] = f2bf(hv[it].z);
      ar[3] = f2bf(hv[it].w);
      ar[WID + 0] = f2bf(av[it].x - hv[it].x);
      ar[WID + 1] = f2bf(av[it].y - hv[it].y);
      ar[WID + 2] = f2bf(av[it].z - hv[it].z);
      ar[WID + 3] = f2bf(av[it].w - hv[it].w);
    }
    __syncthreads();
    f32x4 acc[NTL];
#pragma unroll
    for (int t = 0; t < NTL; t++) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int k = 0; k < KS; k++) {
      bf16x4 a;
      const bf16_t *ap = &As[(w * 16 + c) * ASTR + k * 16 + 4 * g];
#pragma unroll
      for (int jj = 0; jj < 4; jj++) a[jj] = (short)ap[jj];
#pragma unroll
      for (int t = 0; t < NTL; t++) {
        const uint64_t bv = Bf[(k * NTL + t) * 64 + l];
        bf16x4 b;
#pragma unroll
        for (int jj = 0; jj < 4; jj++) b[jj] = (short)(uint16_t)(bv >> (16 * jj));
        acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[t], 0, 0, 0);
      }
    }
    // epilogue: lane holds rows w*16 + 4g + jj, columns t*16 + c; h_l re-read from
    // memory (the tile was just staged, so from cache)
#pragma unroll
    for (int jj = 0; jj < 4; jj++) {
      const int64_t row = row0 + w * 16 + 4 * g + jj;
      const bool ok = row < rows;
      float ud = 0.f;
#pragma unroll
      for (int t = 0; t < NTL; t++) {
        const int col = t * 16 + c;
        const float h = ok ? hin[row * WID + col] : 0.f;
        const float hn = h + fmaxf(acc[t][jj] + bcol[t], 0.f);
        if (!wd) {
          if (ok) hout[row * WID + col] = hn;
        } else {
          ud += hn * wdc[t];
        }
      }
      if (wd) {
#pragma unroll
        for (int o = 8; o > 0; o >>= 1) ud += __shfl_xor(ud, o, 16);
        if (c == 0 && ok) u[row] = ud + bd[0];
      }
    }
  }
}

// The 8 input features of a 2D family with one source field and a Dirichlet value
// per side of the bounding box, computed on the device (row r = sample r / n, node
// r % n): f / 5, asinh(f / 2), side value k * exp(-distance to side k / delta)
// (k = south, east, north, west), x, y.  The host-side definition is identical.
__global__ void k_features_sides(int64_t rows, int64_t n, const double *__restrict__ f,
                                 const double *__restrict__ sides, const float *__restrict__ xy,
                                 const float *__restrict__ dist, double inv_delta, float *__restrict__ x) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= rows) return;
  const int64_t smp = r / n, i = r - smp * n;
  const double fv = f[r];
  float *o = x + r * NFEAT;
  o[0] = (float)(fv / 5.0);
  o[1] = (float)asinh(fv / 2.0);
#pragma unroll
  for (int k = 0; k < 4; k++) o[2 + k] = (float)sides[smp * 4 + k] * (float)exp(-(double)dist[i * 4 + k] * inv_delta);
  o[6] = xy[2 * i];
  o[7] = xy[2 * i + 1];
}

// network input (rows x 16 float) -> bfloat16
__global__ void k_pad_features(int64_t rows, const float *__restrict__ x16, bf16_t *__restrict__ xb) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= rows * KPAD) return;
  xb[q] = f2bf(x16[q]);
}

// ---------------------------------------------------------- context -------
// Channels 8..15 of the network input: two features smoothed on the graph by
// K = 16, 64, 160, 324 steps of y <- (y + W y) / 2.  All samples share the graph, so
// the fields of a launch are kept node-major, Y[i][col] with col = c * ns + s (feature
// c, sample s): a wave takes one node and 64 columns, the node's graph entries are
// wave-uniform and every field load is a coalesced row segment.  One launch per step.
constexpr int SMOOTH_STEPS_H[4] = {16, 64, 160, 324};
constexpr int SMOOTH_SRC[2] = {1, 0};  // asinh(f/2), f/5

// x16[r, 0..7] = x[r, 0..7]
__global__ void k_expand_copy(int64_t rows, const float *__restrict__ x, float *__restrict__ x16) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= rows * KPAD) return;
  const int64_t r = q / KPAD;
  const int k = (int)(q % KPAD);
  if (k < NFEAT) x16[q] = x[r * NFEAT + k];
}

// Y[i][c ns + s] = x16[s n + i][src_c]  (to node-major; row stride Cs >= 2 ns, a multiple
// of 4, padding columns zero), or back: x16[..][slot_c] = Y
template <bool TO_NODE_MAJOR>
__global__ void k_smooth_io(int64_t n, int64_t ns, int64_t Cs, float *__restrict__ x16, float *__restrict__ Y,
                            int snap) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n * Cs) return;
  const int64_t i = q / Cs, col = q - i * Cs;
  if (col >= 2 * ns) {
    if (TO_NODE_MAJOR) Y[q] = 0.f;
    return;
  }
  const int c = (int)(col / ns);
  const int64_t s = col - (int64_t)c * ns;
  float *xr = x16 + (s * n + i) * KPAD;
  if (TO_NODE_MAJOR)
    Y[q] = xr[c == 0 ? SMOOTH_SRC[0] : SMOOTH_SRC[1]];
  else
    xr[NFEAT + 4 * c + snap] = Y[q];
}

// one step: Yo[i][:] = (Yi[i][:] + sum_j w_ij Yi[j][:]) / 2.  A wave takes one node (its
// graph entries are wave-uniform) and walks the row in float4 columns; the neighbour
// rows are loaded four at a time and summed in entry order.
__global__ void __launch_bounds__(256) k_smooth_step(int64_t n, int64_t C4, const int32_t *__restrict__ gptr,
                                                     const int32_t *__restrict__ gcol,
                                                     const float *__restrict__ gw, const float4 *__restrict__ Yi,
                                                     float4 *__restrict__ Yo) {
  const int64_t i = __builtin_amdgcn_readfirstlane(
      (int)(xcd_block(blockIdx.x, gridDim.x) * 4 + (threadIdx.x >> 6)));
  if (i >= n) return;
  const int32_t k0 = gptr[i], k1 = gptr[i + 1];
  for (int64_t c4 = (int64_t)blockIdx.y * 64 + (threadIdx.x & 63); c4 < C4; c4 += (int64_t)gridDim.y * 64) {
    float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
    int32_t k = k0;
    for (; k + 4 <= k1; k += 4) {
      float wk[4];
      float4 v[4];
#pragma unroll
      for (int jj = 0; jj < 4; jj++) {
        wk[jj] = gw[k + jj];
        v[jj] = Yi[(int64_t)gcol[k + jj] * C4 + c4];
      }
#pragma unroll
      for (int jj = 0; jj < 4; jj++) {
        acc.x += wk[jj] * v[jj].x;
        acc.y += wk[jj] * v[jj].y;
        acc.z += wk[jj] * v[jj].z;
        acc.w += wk[jj] * v[jj].w;
      }
    }
    for (; k < k1; k++) {
      const float wv = gw[k];
      const float4 v = Yi[(int64_t)gcol[k] * C4 + c4];
      acc.x += wv * v.x;
      acc.y += wv * v.y;
      acc.z += wv * v.z;
      acc.w += wv * v.w;
    }
    const float4 y = Yi[i * C4 + c4];
    Yo[i * C4 + c4] = make_float4(0.5f * (y.x + acc.x), 0.5f * (y.y + acc.y), 0.5f * (y.z + acc.z),
                                  0.5f * (y.w + acc.w));
  }
}

// u_r = h_L[r] . wd + bd (one wave per row); optional loss: du = 2 (u - t) / count,
// dh = du wd, per-block partials of sum (u-t)^2, sum du h (64) and sum du
__global__ void __launch_bounds__(TB) k_decode(int64_t rows, const float *__restrict__ hf, const float *__restrict__ wd,
                                               const float *__restrict__ bd, float *__restrict__ u,
                                               const float *__restrict__ target, float inv_count,
                                               float *__restrict__ dh, float *__restrict__ part) {
  __shared__ float sh_loss[4], sh_du[4];
  __shared__ float sh_g[4][WID];
  const int w = threadIdx.x >> 6, f = threadIdx.x & 63;
  const int64_t row = (int64_t)blockIdx.x * 4 + w;
  float hv = 0.f, du = 0.f, lo = 0.f;
  if (row < rows) {
    hv = hf[row * WID + f];
    float v = hv * wd[f];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    const float uu = v + bd[0];
    if (f == 0) u[row] = uu;
    if (target) {
      const float d = uu - target[row];
      du = 2.f * d * inv_count;
      lo = d * d;
      dh[row * WID + f] = du * wd[f];
    }
  }
  if (!target) return;
  sh_g[w][f] = du * hv;
  if (f == 0) {
    sh_loss[w] = lo;
    sh_du[w] = du;
  }
  __syncthreads();
  if (w == 0) {
    float gsum = sh_g[0][f] + sh_g[1][f] + sh_g[2][f] + sh_g[3][f];
    float *pb = part + (size_t)blockIdx.x * (WID + 2);
    pb[f] = gsum;
    if (f == 0) {
      pb[WID] = sh_loss[0] + sh_loss[1] + sh_loss[2] + sh_loss[3];
      pb[WID + 1] = sh_du[0] + sh_du[1] + sh_du[2] + sh_du[3];
    }
  }
}

// dz = dh * (gate > 0) -> bfloat16; per-block column sums of dz (bias gradient)
// gate is r_l (float) or, for the encoder, bf(h0) in the first half of cat_0
template <bool GATE_BF>
__global__ void __launch_bounds__(TB) k_gate(int64_t rows, const float *__restrict__ dh, const float *__restrict__ gate,
                                             const bf16_t *__restrict__ gate_bf, bf16_t *__restrict__ dzb,
                                             float *__restrict__ part, int rows_per_block) {
  __shared__ float sh[4][WID];
  const int w = threadIdx.x >> 6, f = threadIdx.x & 63;
  const int64_t r0 = (int64_t)blockIdx.x * rows_per_block;
  const int64_t r1 = std::min<int64_t>(rows, r0 + rows_per_block);
  float acc = 0.f;
  for (int64_t row = r0 + w; row < r1; row += 4) {
    const bool on = GATE_BF ? (bf2f(gate_bf[row * CW + f]) > 0.f) : (gate[row * WID + f] > 0.f);
    const float dz = on ? dh[row * WID + f] : 0.f;
    dzb[row * WID + f] = f2bf(dz);
    acc += dz;
  }
  sh[w][f] = acc;
  __syncthreads();
  if (w == 0) part[(size_t)blockIdx.x * WID + f] = sh[0][f] + sh[1][f] + sh[2][f] + sh[3][f];
}

// Adam on the float32 parameters (t = step count, bias-corrected)
__global__ void k_adam(int64_t n, float *__restrict__ theta, const float *__restrict__ grad, float *__restrict__ m,
                       float *__restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const float g = grad[k];
  const float mk = b1 * m[k] + (1.f - b1) * g;
  const float vk = b2 * v[k] + (1.f - b2) * g * g;
  m[k] = mk;
  v[k] = vk;
  theta[k] -= lr * (mk / c1) / (sqrtf(vk / c2) + eps);
}

// bfloat16 copies: W (K x N) as is and transposed
__global__ void k_cast(int K, int N, const float *__restrict__ W, bf16_t *__restrict__ Wb, bf16_t *__restrict__ WTb) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= K * N) return;
  const int k = q / N, j = q % N;
  const bf16_t b = f2bf(W[q]);
  Wb[q] = b;
  if (WTb) WTb[j * K + k] = b;
}

__global__ void k_fill_f(int64_t n, float *p, float v) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) p[k] = v;
}

// u (float, rows) -> out (double)
__global__ void k_to_double(int64_t rows, const float *__restrict__ u, double *__restrict__ out) {
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r < rows) out[r] = (double)u[r];
}

}  // namespace

struct vfvm_mpnn {
  vfvm_ctx *ctx = nullptr;
  int nlayers = 4;
  int64_t n = 0;  // nodes per sample
  int64_t np = 0;
  // graph
  DevBuf<int32_t> gptr, gcol;
  DevBuf<float> gw, gwt;
  // parameters
  DevBuf<float> theta, grad, adam_m, adam_v;
  int64_t adam_t = 0;
  DevBuf<bf16_t> Web, Wb, WTb;  // encoder; layers (L x 128 x 64) and transposes (L x 64 x 128)
  // activations for a batch of `cap` rows
  int64_t cap = 0;
  DevBuf<float> x, x16, hf, rbuf, dh, da, u, target, part, sm0, sm1;
  DevBuf<bf16_t> xb, cat, dzb;
  // device features of the Dirichlet-sides family (2D): per node x, y and the distances
  // to the bounding-box sides (south, east, north, west); per-sample inputs
  DevBuf<float> xy, dist;
  DevBuf<double> fin, sides;
  size_t off_We = 0, off_be = 0, off_wd = 0, off_bd = 0;
  std::vector<size_t> off_W, off_b;
};

namespace {

template <class F>
int mguard(vfvm_ctx *ctx, F &&f) {
  try {
    return f();
  } catch (const Error &e) {
    if (ctx) ctx->err = e.what();
    return e.code;
  } catch (const std::exception &e) {
    if (ctx) ctx->err = e.what();
    return VFVM_ERR_HIP;
  }
}

inline unsigned nb(int64_t n, int t) { return (unsigned)((n + t - 1) / t); }

void refresh_bf16(vfvm_mpnn &m) {
  hipStream_t st = m.ctx->stream;
  hipLaunchKernelGGL(k_cast, dim3(nb(KPAD * WID, 256)), dim3(256), 0, st, KPAD, WID, (const float *)m.theta.p + m.off_We,
                     m.Web.p, (bf16_t *)nullptr);
  for (int l = 0; l < m.nlayers; l++)
    hipLaunchKernelGGL(k_cast, dim3(nb(CW * WID, 256)), dim3(256), 0, st, CW, WID,
                       (const float *)m.theta.p + m.off_W[l], m.Wb.p + (size_t)l * CW * WID,
                       m.WTb.p + (size_t)l * CW * WID);
  VFVM_HIP_CHECK(hipGetLastError());
}

// samples per inference launch sequence: at most MAX_INFER_ROWS rows, the samples split
// into chunks of equal size (a short last chunk would pay every launch of the smoothing
// and the layers for a few samples)
constexpr int64_t MAX_INFER_ROWS = int64_t(1) << 24;
inline int64_t sample_chunk(int64_t nsamples, int64_t n) {
  const int64_t most = std::max<int64_t>(1, MAX_INFER_ROWS / n);
  const int64_t nchunk = (nsamples + most - 1) / most;
  return std::max<int64_t>(1, (nsamples + nchunk - 1) / nchunk);
}

constexpr int64_t WG_ROWS = 4096;  // rows per block of the weight-gradient kernels
constexpr int GATE_ROWS = 256;     // rows per block of the gate kernel

void ensure_cap(vfvm_mpnn &m, int64_t rows) {
  if (rows <= m.cap) return;
  const int L = m.nlayers;
  m.x.alloc((size_t)rows * NFEAT);
  m.x16.alloc((size_t)rows * KPAD);
  m.xb.alloc((size_t)rows * KPAD);
  m.hf.alloc((size_t)rows * WID);
  m.rbuf.alloc((size_t)rows * WID * L);
  m.dh.alloc((size_t)rows * WID);
  m.da.alloc((size_t)rows * WID);
  m.u.alloc((size_t)rows);
  m.target.alloc((size_t)rows);
  m.cat.alloc((size_t)rows * CW * (L + 1));
  m.dzb.alloc((size_t)rows * WID);
  const int64_t p1 = nb(rows, 4) * (int64_t)(WID + 2);
  const int64_t p2 = (int64_t)nb(rows, WG_ROWS) * CW * WID;
  const int64_t p3 = (int64_t)nb(rows, GATE_ROWS) * WID;
  m.part.alloc((size_t)std::max({p1, p2, p3}));
  m.cap = rows;
}

// network input of `ns` samples: x16 = [x (8 features), smoothed context (8 channels)]
void expand(vfvm_mpnn &m, const float *x, float *x16, int64_t ns) {
  hipStream_t st = m.ctx->stream;
  const int64_t rows = ns * m.n, Cs = (2 * ns + 3) / 4 * 4, C4 = Cs / 4;
  hipLaunchKernelGGL(k_expand_copy, dim3(nb(rows * KPAD, 256)), dim3(256), 0, st, rows, x, x16);
  m.sm0.alloc((size_t)std::max<int64_t>(m.n * Cs, (int64_t)m.sm0.n));
  m.sm1.alloc((size_t)std::max<int64_t>(m.n * Cs, (int64_t)m.sm1.n));
  hipLaunchKernelGGL((k_smooth_io<true>), dim3(nb(m.n * Cs, 256)), dim3(256), 0, st, m.n, ns, Cs, x16, m.sm0.p, 0);
  float *a = m.sm0.p, *b = m.sm1.p;
  const unsigned gy = (unsigned)std::max<int64_t>(1, std::min<int64_t>(C4 / 256, 64));
  int snap = 0;
  for (int step = 1; step <= SMOOTH_STEPS_H[3]; step++) {
    hipLaunchKernelGGL(k_smooth_step, dim3((nb(m.n, 4) + 7) / 8 * 8, gy), dim3(256), 0, st, m.n, C4, (const int32_t *)m.gptr.p,
                       (const int32_t *)m.gcol.p, (const float *)m.gw.p, (const float4 *)a, (float4 *)b);
    std::swap(a, b);
    if (step == SMOOTH_STEPS_H[snap]) {
      hipLaunchKernelGGL((k_smooth_io<false>), dim3(nb(m.n * Cs, 256)), dim3(256), 0, st, m.n, ns, Cs, x16, a,
                         snap);
      snap++;
    }
  }
  VFVM_HIP_CHECK(hipGetLastError());
}

// forward pass on `rows` rows whose network input is in m.x16; keeps what the backward pass needs
void forward(vfvm_mpnn &m, int64_t rows) {
  hipStream_t st = m.ctx->stream;
  const int L = m.nlayers;
  hipLaunchKernelGGL(k_pad_features, dim3(nb(rows * KPAD, 256)), dim3(256), 0, st, rows, (const float *)m.x16.p,
                     m.xb.p);
  GemmEpi e{};
  e.bias = m.theta.p + m.off_be;
  e.hf = m.hf.p;
  e.cat = m.cat.p;  // cat_0 first half = bf(h0)
  hipLaunchKernelGGL((k_gemm<KPAD, WID, EPI_ENC>), dim3(nb(rows, 64)), dim3(TB), 0, st, rows, (const bf16_t *)m.xb.p,
                     (const bf16_t *)m.Web.p, e);
  for (int l = 0; l < L; l++) {
    bf16_t *cl = m.cat.p + (size_t)l * rows * CW;
    hipLaunchKernelGGL(k_aggregate, dim3(nb(rows, 4)), dim3(TB), 0, st, rows, m.n, (const int32_t *)m.gptr.p,
                       (const int32_t *)m.gcol.p, (const float *)m.gw.p, (const float *)m.hf.p, cl);
    GemmEpi r{};
    r.bias = m.theta.p + m.off_b[l];
    r.hf = m.hf.p;
    r.rbuf = m.rbuf.p + (size_t)l * rows * WID;
    r.cat = m.cat.p + (size_t)(l + 1) * rows * CW;  // next layer's h half (cat_L unused by the decoder)
    hipLaunchKernelGGL((k_gemm<CW, WID, EPI_RES>), dim3(nb(rows, 64)), dim3(TB), 0, st, rows, (const bf16_t *)cl,
                       (const bf16_t *)(m.Wb.p + (size_t)l * CW * WID), r);
  }
  VFVM_HIP_CHECK(hipGetLastError());
}

// inference: encoder, then the fused layers ping-ponging hf / dh; u = network output
void forward_infer(vfvm_mpnn &m, int64_t rows) {
  hipStream_t st = m.ctx->stream;
  const int L = m.nlayers;
  hipLaunchKernelGGL(k_pad_features, dim3(nb(rows * KPAD, 256)), dim3(256), 0, st, rows, (const float *)m.x16.p,
                     m.xb.p);
  GemmEpi e{};
  e.bias = m.theta.p + m.off_be;
  e.hf = m.hf.p;
  hipLaunchKernelGGL((k_gemm<KPAD, WID, EPI_ENC>), dim3(nb(rows, 64)), dim3(TB), 0, st, rows, (const bf16_t *)m.xb.p,
                     (const bf16_t *)m.Web.p, e);
  const unsigned grid = (unsigned)std::min<int64_t>((nb(rows, 64) + 7) / 8 * 8, 2048);  // a multiple of 8
  float *hin = m.hf.p, *hout = m.dh.p;
  for (int l = 0; l < L; l++) {
    const bool last = l == L - 1;
    hipLaunchKernelGGL(k_layer_infer, dim3(grid), dim3(TB), 0, st, rows, m.n, (const int32_t *)m.gptr.p,
                       (const int32_t *)m.gcol.p, (const float *)m.gw.p, (const float *)hin,
                       (const bf16_t *)(m.Wb.p + (size_t)l * CW * WID), (const float *)m.theta.p + m.off_b[l],
                       last ? (float *)nullptr : hout, last ? (const float *)m.theta.p + m.off_wd : nullptr,
                       last ? (const float *)m.theta.p + m.off_bd : nullptr, last ? m.u.p : (float *)nullptr);
    std::swap(hin, hout);
  }
  VFVM_HIP_CHECK(hipGetLastError());
}

// loss and full gradient (into m.grad) for `rows` rows with targets in m.target
float backward(vfvm_mpnn &m, int64_t rows) {
  hipStream_t st = m.ctx->stream;
  const int L = m.nlayers;
  const unsigned dblk = nb(rows, 4);
  hipLaunchKernelGGL(k_decode, dim3(dblk), dim3(TB), 0, st, rows, (const float *)m.hf.p,
                     (const float *)m.theta.p + m.off_wd, (const float *)m.theta.p + m.off_bd, m.u.p,
                     (const float *)m.target.p, 1.0f / (float)rows, m.dh.p, m.part.p);
  // decoder gradient and loss from the per-block partials (64 + 2 per block)
  std::vector<float> hpart((size_t)dblk * (WID + 2));
  VFVM_HIP_CHECK(hipMemcpyAsync(hpart.data(), m.part.p, hpart.size() * sizeof(float), hipMemcpyDeviceToHost, st));
  VFVM_HIP_CHECK(hipStreamSynchronize(st));
  std::vector<double> gwd(WID, 0.0);
  double loss = 0.0, gbd = 0.0;
  for (unsigned b = 0; b < dblk; b++) {
    for (int f = 0; f < WID; f++) gwd[f] += hpart[(size_t)b * (WID + 2) + f];
    loss += hpart[(size_t)b * (WID + 2) + WID];
    gbd += hpart[(size_t)b * (WID + 2) + WID + 1];
  }
  std::vector<float> gdec(WID + 1);
  for (int f = 0; f < WID; f++) gdec[f] = (float)gwd[f];
  gdec[WID] = (float)gbd;
  VFVM_HIP_CHECK(hipMemcpyAsync(m.grad.p + m.off_wd, gdec.data(), (WID + 1) * sizeof(float), hipMemcpyHostToDevice, st));
  VFVM_HIP_CHECK(hipStreamSynchronize(st));  // gdec is a local
  const unsigned wg_blocks = nb(rows, WG_ROWS);
  const unsigned gate_blocks = nb(rows, GATE_ROWS);
  for (int l = L - 1; l >= 0; l--) {
    const bf16_t *cl = m.cat.p + (size_t)l * rows * CW;
    hipLaunchKernelGGL((k_gate<false>), dim3(gate_blocks), dim3(TB), 0, st, rows, (const float *)m.dh.p,
                       (const float *)m.rbuf.p + (size_t)l * rows * WID, (const bf16_t *)nullptr, m.dzb.p, m.part.p,
                       GATE_ROWS);
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(WID), 0, st, (int)gate_blocks, WID, (const float *)m.part.p,
                       m.grad.p + m.off_b[l]);
    hipLaunchKernelGGL((k_wgrad<CW>), dim3(wg_blocks), dim3(TB), 0, st, rows, WG_ROWS, cl, (const bf16_t *)m.dzb.p,
                       m.part.p);
    hipLaunchKernelGGL(k_sum_partials, dim3(nb(CW * WID, 256)), dim3(256), 0, st, (int)wg_blocks, CW * WID,
                       (const float *)m.part.p, m.grad.p + m.off_W[l]);
    GemmEpi e{};
    e.dh = m.dh.p;
    e.da = m.da.p;
    hipLaunchKernelGGL((k_gemm<WID, CW, EPI_DCAT>), dim3(nb(rows, 64)), dim3(TB), 0, st, rows, (const bf16_t *)m.dzb.p,
                       (const bf16_t *)(m.WTb.p + (size_t)l * CW * WID), e);
    hipLaunchKernelGGL(k_aggregate_t, dim3(nb(rows, 4)), dim3(TB), 0, st, rows, m.n, (const int32_t *)m.gptr.p,
                       (const int32_t *)m.gcol.p, (const float *)m.gwt.p, (const float *)m.da.p, m.dh.p);
  }
  // encoder: gate from bf(h0) in cat_0
  hipLaunchKernelGGL((k_gate<true>), dim3(gate_blocks), dim3(TB), 0, st, rows, (const float *)m.dh.p,
                     (const float *)nullptr, (const bf16_t *)m.cat.p, m.dzb.p, m.part.p, GATE_ROWS);
  hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(WID), 0, st, (int)gate_blocks, WID, (const float *)m.part.p,
                     m.grad.p + m.off_be);
  hipLaunchKernelGGL((k_wgrad<KPAD>), dim3(wg_blocks), dim3(TB), 0, st, rows, WG_ROWS, (const bf16_t *)m.xb.p,
                     (const bf16_t *)m.dzb.p, m.part.p);
  hipLaunchKernelGGL(k_sum_partials, dim3(nb(KPAD * WID, 256)), dim3(256), 0, st, (int)wg_blocks, KPAD * WID,
                     (const float *)m.part.p, m.grad.p + m.off_We);
  VFVM_HIP_CHECK(hipGetLastError());
  return (float)(loss / (double)rows);
}

// inference with the Dirichlet-sides features computed on the device from the sources f
// (nsamples x n) and the side values (nsamples x 4); host or device arrays
int infer_sides_impl(vfvm_mpnn *m, int64_t nsamples, const double *f, const double *sides, double eps, double *out,
                     bool on_device) {
  return mguard(m->ctx, [&]() {
    if (!m->xy.p) throw Error(VFVM_ERR_ARG, "side features need a 2D system");
    hipStream_t st = m->ctx->stream;
    const int64_t chunk = sample_chunk(nsamples, m->n);
    const double inv_delta = 1.0 / std::sqrt(eps / 2.0);
    std::vector<float> u;
    for (int64_t s0 = 0; s0 < nsamples; s0 += chunk) {
      const int64_t ns = std::min(chunk, nsamples - s0), rows = ns * m->n;
      ensure_cap(*m, rows);
      const double *fp = f + s0 * m->n, *sp = sides + s0 * 4;
      if (!on_device) {
        m->fin.upload(fp, (size_t)rows, st);
        m->sides.upload(sp, (size_t)ns * 4, st);
        fp = m->fin.p;
        sp = m->sides.p;
      }
      hipLaunchKernelGGL(k_features_sides, dim3(nb(rows, 256)), dim3(256), 0, st, rows, m->n, fp, sp,
                         (const float *)m->xy.p, (const float *)m->dist.p, inv_delta, m->x.p);
      expand(*m, m->x.p, m->x16.p, ns);
      forward_infer(*m, rows);
      if (on_device) {
        hipLaunchKernelGGL(k_to_double, dim3(nb(rows, 256)), dim3(256), 0, st, rows, (const float *)m->u.p,
                           out + s0 * m->n);
        VFVM_HIP_CHECK(hipGetLastError());
      } else {
        u.resize((size_t)rows);
        m->u.download(u.data(), (size_t)rows, st);
        VFVM_HIP_CHECK(hipStreamSynchronize(st));
        for (int64_t k = 0; k < rows; k++) out[s0 * m->n + k] = (double)u[k];
      }
    }
    VFVM_HIP_CHECK(hipStreamSynchronize(st));
    return VFVM_OK;
  });
}

}  // namespace

extern "C" {

int vfvm_mpnn_create(vfvm_sys *sys, int nlayers, int width, uint64_t seed, vfvm_mpnn **out) {
  if (!sys || !out) return VFVM_ERR_ARG;
  return mguard(sys->ctx, [&]() {
    if (width != WID) throw Error(VFVM_ERR_ARG, "the network width is fixed at 64");
    if (nlayers < 1 || nlayers > 16) throw Error(VFVM_ERR_ARG, "nlayers must be 1..16");
    std::unique_ptr<vfvm_mpnn> m(new vfvm_mpnn());
    m->ctx = sys->ctx;
    m->nlayers = nlayers;
    m->n = sys->nnodes;
    // normalised graph weights from the system's node graph (sigma/h factors)
    sync_host_geometry(*sys);
    const HostSetup &h = sys->hs;
    std::vector<int32_t> ptr((size_t)m->n + 1, 0), col;
    std::vector<float> w, wt;
    std::vector<double> rowsum((size_t)m->n, 0.0);
    for (int64_t i = 0; i < m->n; i++)
      for (int32_t k = h.adj_ptr[i]; k < h.adj_ptr[i + 1]; k++)
        if (h.adj_col[k] != i) rowsum[i] += h.adj_fac[k];
    for (int64_t i = 0; i < m->n; i++) {
      for (int32_t k = h.adj_ptr[i]; k < h.adj_ptr[i + 1]; k++) {
        const int32_t j = h.adj_col[k];
        if (j == i) continue;
        col.push_back(j);
        w.push_back(rowsum[i] > 0 ? (float)(h.adj_fac[k] / rowsum[i]) : 0.f);
        wt.push_back(rowsum[j] > 0 ? (float)(h.adj_fac[k] / rowsum[j]) : 0.f);  // w_ji on i's entry
      }
      ptr[i + 1] = (int32_t)col.size();
    }
    hipStream_t st = m->ctx->stream;
    if (sys->dim == 2) {
      const double *X = sys->coord.data();
      double lo[2] = {X[0], X[1]}, hi[2] = {X[0], X[1]};
      for (int64_t i = 0; i < m->n; i++)
        for (int d = 0; d < 2; d++) {
          lo[d] = std::min(lo[d], X[2 * i + d]);
          hi[d] = std::max(hi[d], X[2 * i + d]);
        }
      std::vector<float> xy((size_t)m->n * 2), dist((size_t)m->n * 4);
      for (int64_t i = 0; i < m->n; i++) {
        const double x = X[2 * i], y = X[2 * i + 1];
        xy[2 * i] = (float)x;
        xy[2 * i + 1] = (float)y;
        dist[4 * i + 0] = (float)(y - lo[1]);
        dist[4 * i + 1] = (float)(hi[0] - x);
        dist[4 * i + 2] = (float)(hi[1] - y);
        dist[4 * i + 3] = (float)(x - lo[0]);
      }
      m->xy.upload(xy, st);
      m->dist.upload(dist, st);
    }
    m->gptr.upload(ptr, st);
    m->gcol.upload(col.empty() ? std::vector<int32_t>{0} : col, st);
    m->gw.upload(w.empty() ? std::vector<float>{0.f} : w, st);
    m->gwt.upload(wt.empty() ? std::vector<float>{0.f} : wt, st);
    // parameter layout and He-uniform initialisation
    size_t p = 0;
    m->off_We = p;
    p += KPAD * WID;
    m->off_be = p;
    p += WID;
    for (int l = 0; l < nlayers; l++) {
      m->off_W.push_back(p);
      p += CW * WID;
      m->off_b.push_back(p);
      p += WID;
    }
    m->off_wd = p;
    p += WID;
    m->off_bd = p;
    p += 1;
    m->np = (int64_t)p;
    std::vector<float> th(p, 0.f);
    std::mt19937_64 rng(seed);
    std::uniform_real_distribution<float> U(-1.f, 1.f);
    for (int k = 0; k < KPAD * WID; k++)
      th[m->off_We + k] = U(rng) * std::sqrt(6.f / KPAD);
    for (int l = 0; l < nlayers; l++)
      for (int k = 0; k < CW * WID; k++) th[m->off_W[l] + k] = U(rng) * std::sqrt(6.f / CW) * 0.5f;
    for (int k = 0; k < WID; k++) th[m->off_wd + k] = U(rng) * std::sqrt(3.f / WID);
    m->theta.upload(th, st);
    m->grad.alloc(p);
    m->adam_m.alloc(p);
    m->adam_v.alloc(p);
    hipLaunchKernelGGL(k_fill_f, dim3(nb(p, 256)), dim3(256), 0, st, (int64_t)p, m->adam_m.p, 0.f);
    hipLaunchKernelGGL(k_fill_f, dim3(nb(p, 256)), dim3(256), 0, st, (int64_t)p, m->adam_v.p, 0.f);
    hipLaunchKernelGGL(k_fill_f, dim3(nb(p, 256)), dim3(256), 0, st, (int64_t)p, m->grad.p, 0.f);
    m->Web.alloc(KPAD * WID);
    m->Wb.alloc((size_t)nlayers * CW * WID);
    m->WTb.alloc((size_t)nlayers * CW * WID);
    refresh_bf16(*m);
    VFVM_HIP_CHECK(hipStreamSynchronize(st));
    *out = m.release();
    return VFVM_OK;
  });
}

int vfvm_mpnn_destroy(vfvm_mpnn *m) {
  if (!m) return VFVM_OK;
  (void)hipStreamSynchronize(m->ctx->stream);
  delete m;
  return VFVM_OK;
}

int vfvm_mpnn_params(vfvm_mpnn *m, int64_t *nparams, float *params) {
  if (!m) return VFVM_ERR_ARG;
  return mguard(m->ctx, [&]() {
    if (nparams) *nparams = m->np;
    if (params) {
      m->theta.download(params, (size_t)m->np, m->ctx->stream);
      VFVM_HIP_CHECK(hipStreamSynchronize(m->ctx->stream));
    }
    return VFVM_OK;
  });
}

int vfvm_mpnn_set_params(vfvm_mpnn *m, const float *params) {
  if (!m || !params) return VFVM_ERR_ARG;
  return mguard(m->ctx, [&]() {
    m->theta.upload(params, (size_t)m->np, m->ctx->stream);
    refresh_bf16(*m);
    VFVM_HIP_CHECK(hipStreamSynchronize(m->ctx->stream));
    return VFVM_OK;
  });
}

int vfvm_mpnn_infer(vfvm_mpnn *m, int64_t nsamples, const float *features, double *out) {
  if (!m || !features || !out || nsamples < 1) return VFVM_ERR_ARG;
  return mguard(m->ctx, [&]() {
    hipStream_t st = m->ctx->stream;
    const int64_t chunk = sample_chunk(nsamples, m->n);
    std::vector<float> u;
    for (int64_t s0 = 0; s0 < nsamples; s0 += chunk) {
      const int64_t ns = std::min(chunk, nsamples - s0), rows = ns * m->n;
      ensure_cap(*m, rows);
      m->x.upload(features + s0 * m->n * NFEAT, (size_t)rows * NFEAT, st);
      expand(*m, m->x.p, m->x16.p, ns);
      forward_infer(*m, rows);
      u.resize((size_t)rows);
      m->u.download(u.data(), (size_t)rows, st);
      VFVM_HIP_CHECK(hipStreamSynchronize(st));
      for (int64_t k = 0; k < rows; k++) out[s0 * m->n + k] = (double)u[k];
    }
    return VFVM_OK;
  });
}

int vfvm_mpnn_infer_sides(vfvm_mpnn *m, int64_t nsamples, const double *f, const double *sides, double eps,
                          double *out) {
  if (!m || !f || !sides || !out || nsamples < 1 || !(eps > 0)) return VFVM_ERR_ARG;
  return infer_sides_impl(m, nsamples, f, sides, eps, out, false);
}

int vfvm_mpnn_infer_sides_device(vfvm_mpnn *m, int64_t nsamples, const double *f_dev, const double *sides_dev,
                                 double eps, double *out_dev) {
  if (!m || !f_dev || !sides_dev || !out_dev || nsamples < 1 || !(eps > 0)) return VFVM_ERR_ARG;
  return infer_sides_impl(m, nsamples, f_dev, sides_dev, eps, out_dev, true);
}

// average device time (ms) of one inference hidden layer (k_layer_infer, not the last)
// over `rows` rows of the current hidden state
int vfvm_mpnn_time_layer(vfvm_mpnn *m, int64_t rows, int reps, double *avg_ms) {
  if (!m || rows < 1 || reps < 1 || !avg_ms) return VFVM_ERR_ARG;
  return mguard(m->ctx, [&]() {
    if (rows % m->n != 0) throw Error(VFVM_ERR_ARG, "rows must be a whole number of samples");
    hipStream_t st = m->ctx->stream;
    ensure_cap(*m, rows);
    const unsigned grid = (unsigned)std::min<int64_t>((nb(rows, 64) + 7) / 8 * 8, 2048);
    auto once = [&]() {
      hipLaunchKernelGGL(k_layer_infer, dim3(grid), dim3(TB), 0, st, rows, m->n, (const int32_t *)m->gptr.p,
                         (const int32_t *)m->gcol.p, (const float *)m->gw.p, (const float *)m->hf.p,
                         (const bf16_t *)m->Wb.p, (const float *)m->theta.p + m->off_b[0], m->dh.p,
                         (const float *)nullptr, (const float *)nullptr, (float *)nullptr);
    };
    hipLaunchKernelGGL(k_fill_f, dim3(nb(rows * WID, 256)), dim3(256), 0, st, rows * WID, m->hf.p, 0.5f);
    once();
    hipEvent_t e0, e1;
    VFVM_HIP_CHECK(hipEventCreate(&e0));
    VFVM_HIP_CHECK(hipEventCreate(&e1));
    VFVM_HIP_CHECK(hipEventRecord(e0, st));
    for (int k = 0; k < reps; k++) once();
    VFVM_HIP_CHECK(hipEventRecord(e1, st));
    VFVM_HIP_CHECK(hipEventSynchronize(e1));
    float ms = 0;
    VFVM_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *avg_ms = (double)ms / reps;
    return VFVM_OK;
  });
}

int vfvm_mpnn_loss_grad(vfvm_mpnn *m, int64_t nsamples, const float *features, const float *targets, float *loss,
                        float *grad) {
  if (!m || !features || !targets || nsamples < 1) return VFVM_ERR_ARG;
  return mguard(m->ctx, [&]() {
    hipStream_t st = m->ctx->stream;
    const int64_t rows = nsamples * m->n;
    ensure_cap(*m, rows);
    m->x.upload(features, (size_t)rows * NFEAT, st);
    m->target.upload(targets, (size_t)rows, st);
    expand(*m, m->x.p, m->x16.p, nsamples);
    forward(*m, rows);
    const float lv = backward(*m, rows);
    if (loss) *loss = lv;
    if (grad) {
      m->grad.download(grad, (size_t)m->np, st);
      VFVM_HIP_CHECK(hipStreamSynchronize(st));
    }
    return VFVM_OK;
  });
}

int vfvm_mpnn_train(vfvm_mpnn *m, int64_t nsamples, const float *features, const float *targets, int64_t epochs,
                    int64_t batch, float lr, float *final_loss) {
  if (!m || !features || !targets || nsamples < 1 || batch < 1 || epochs < 1) return VFVM_ERR_ARG;
  return mguard(m->ctx, [&]() {
    hipStream_t st = m->ctx->stream;
    const int64_t n = m->n;
    // the whole training set (with its smoothed context channels) stays resident on the device
    DevBuf<float> X, X16, T;
    X.upload(features, (size_t)nsamples * n * NFEAT, st);
    X16.alloc((size_t)nsamples * n * KPAD);
    const int64_t ech = std::max<int64_t>(1, (int64_t(1) << 22) / n);
    for (int64_t s0 = 0; s0 < nsamples; s0 += ech)
      expand(*m, X.p + s0 * n * NFEAT, X16.p + s0 * n * KPAD, std::min(ech, nsamples - s0));
    X.release();
    T.upload(targets, (size_t)nsamples * n, st);
    const int64_t B = std::min(batch, nsamples);
    ensure_cap(*m, B * n);
    std::vector<int64_t> order((size_t)nsamples);
    for (int64_t k = 0; k < nsamples; k++) order[k] = k;
    std::mt19937_64 rng(1234567u + (uint64_t)m->adam_t);
    double last = 0.0;
    for (int64_t ep = 0; ep < epochs; ep++) {
      std::shuffle(order.begin(), order.end(), rng);
      double acc = 0.0;
      int64_t nbat = 0;
      for (int64_t b0 = 0; b0 + B <= nsamples; b0 += B) {
        for (int64_t q = 0; q < B; q++) {
          const int64_t sidx = order[b0 + q];
          VFVM_HIP_CHECK(hipMemcpyAsync(m->x16.p + q * n * KPAD, X16.p + sidx * n * KPAD, sizeof(float) * n * KPAD,
                                        hipMemcpyDeviceToDevice, st));
          VFVM_HIP_CHECK(hipMemcpyAsync(m->target.p + q * n, T.p + sidx * n, sizeof(float) * n,
                                        hipMemcpyDeviceToDevice, st));
        }
        forward(*m, B * n);
        acc += backward(*m, B * n);
        nbat++;
        m->adam_t++;
        const float b1 = 0.9f, b2 = 0.999f;
        const float c1 = 1.f - std::pow(b1, (float)m->adam_t), c2 = 1.f - std::pow(b2, (float)m->adam_t);
        hipLaunchKernelGGL(k_adam, dim3(nb(m->np, 256)), dim3(256), 0, st, m->np, m->theta.p, (const float *)m->grad.p,
                           m->adam_m.p, m->adam_v.p, lr, b1, b2, 1e-8f, c1, c2);
        refresh_bf16(*m);
      }
      last = nbat ? acc / nbat : 0.0;
    }
    VFVM_HIP_CHECK(hipStreamSynchronize(st));
    if (final_loss) *final_loss = (float)last;
    return VFVM_OK;
  });
}

}  // extern "C"
