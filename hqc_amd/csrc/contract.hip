// AO x grid contractions on MFMA: the two per-iteration hot kernels.
//
//   density:  rho(r) = sum_pq D_pq phi_p(r) phi_q(r)          (solver.py:690)
//   vmat:     V_pq   = dV sum_r phi_p(r) v(r) phi_q(r)        (solver.py:693)
//
// MFMA 16x16x4 operand layout (64 lanes, one value per lane for A and B,
// four per lane for C):  A[i][k] in lane i + 16k, B[k][j] in lane j + 16k.
// The accumulator layout differs between the two instructions (measured on
// gfx950):  f32: C[i][j] in lane j + 16*(i/4), element i%4;
//           f64: C[i][j] in lane j + 16*(i%4), element i/4.
#include "common.h"

__device__ __forceinline__ f64x4 mfma16(double a, double b, f64x4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f32x4 mfma16(float a, float b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// Row (0..15) inside a 16x16 accumulator tile held by lane group kq = lane/16,
// element e.
template <typename T>
__device__ __forceinline__ int acc_row(int kq, int e);
template <>
__device__ __forceinline__ int acc_row<double>(int kq, int e) {
  return e * 4 + kq;
}
template <>
__device__ __forceinline__ int acc_row<float>(int kq, int e) {
  return kq * 4 + e;
}

template <typename T>
struct vec4_of;
template <>
struct vec4_of<double> {
  typedef f64x4 type;
};
template <>
struct vec4_of<float> {
  typedef f32x4 type;
};

// ---------------------------------------------------------------------------
// Density.  Block = 4 waves, 64 grid columns (16 per wave).  For every group of
// NTG row tiles the wave runs the full K loop over AOs: X tile = D(rows, :) *
// phi(:, cols) with D read through L2 (D is symmetric so its row-major rows
// give the A operand coalesced); the epilogue forms sum_p phi_p X_p per column.
// ---------------------------------------------------------------------------
template <typename T, int NTG>
__global__ __launch_bounds__(256) void k_density(const T* __restrict__ phi,
                                                 const double* __restrict__ D, int nao_pad,
                                                 int ld, double* __restrict__ rho,
                                                 double* __restrict__ X) {
  typedef typename vec4_of<T>::type V4;
  const int b = blockIdx.y;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int l16 = lane & 15, kq = lane >> 4;
  const long col = (long)blockIdx.x * 64 + w * 16 + l16;
  const T* P = phi + (long)b * nao_pad * ld;
  const double* Db = D + (long)b * nao_pad * nao_pad;
  const int NT = nao_pad >> 4, KS = nao_pad >> 2;
  double rsum = 0.0;
  for (int rt0 = 0; rt0 < NT; rt0 += NTG) {
    V4 acc[NTG];
#pragma unroll
    for (int i = 0; i < NTG; ++i) acc[i] = V4{0, 0, 0, 0};
    for (int ks = 0; ks < KS; ++ks) {
      const int k = ks * 4 + kq;
      const T bv = P[(long)k * ld + col];
      const double* Drow = Db + (long)k * nao_pad + l16;
#pragma unroll
      for (int i = 0; i < NTG; ++i) {
        if (rt0 + i < NT) {
          const T av = (T)Drow[(rt0 + i) * 16];
          acc[i] = mfma16(av, bv, acc[i]);
        }
      }
    }
#pragma unroll
    for (int i = 0; i < NTG; ++i) {
      if (rt0 + i < NT) {
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const long row = (long)(rt0 + i) * 16 + acc_row<T>(kq, e);
          const double xv = (double)acc[i][e];
          rsum += xv * (double)P[row * ld + col];
          if (X) X[((long)b * nao_pad + row) * ld + col] = xv;
        }
      }
    }
  }
  rsum += __shfl_xor(rsum, 16, 64);
  rsum += __shfl_xor(rsum, 32, 64);
  if (kq == 0) rho[(long)b * ld + col] = rsum;
}

template <typename T>
static int launch_density(const T* phi, const double* D, int B, int nao_pad, int ld,
                          double* rho, double* X, void* stream) {
  if (B < 0 || nao_pad <= 0 || (nao_pad & 15) || ld <= 0 || (ld & 63)) return HQC_E_ARG;
  if (B == 0) return HQC_OK;
  hipLaunchKernelGGL((k_density<T, 8>), dim3(ld / 64, B), dim3(256), 0, (hipStream_t)stream,
                     phi, D, nao_pad, ld, rho, X);
  HQC_CHECK_LAUNCH();
  return HQC_OK;
}

extern "C" int hqc_ao_density_f64(const double* phi, const double* D, int B, int nao_pad,
                                  int ld, double* rho, double* X, void* stream) {
  return launch_density<double>(phi, D, B, nao_pad, ld, rho, X, stream);
}

extern "C" int hqc_ao_density_f32(const float* phi, const double* D, int B, int nao_pad,
                                  int ld, double* rho, void* stream) {
  return launch_density<float>(phi, D, B, nao_pad, ld, rho, nullptr, stream);
}

// ---------------------------------------------------------------------------
// Potential matrix.  Output tiles (rt <= ct, upper triangle) are split into
// groups of 4*TPW (TPW per wave).  Block (wg, group, b) walks grid chunks of KC
// points wg, wg+nwg, ...; each chunk of phi (all rows) and v is staged in LDS
// (row stride KC+1 to spread banks), then every wave runs KC/4 MFMA steps on
// its tiles.  Per-block partial tiles go to `part`; k_vmat_reduce sums them in
// a fixed order and writes both triangles scaled by dV.
// ---------------------------------------------------------------------------
#define VMAT_TPW 8
#define VMAT_TARGET_BLOCKS 512

__device__ __forceinline__ void tile_rc(int t, int NT, int& rt, int& ct) {
  int r = 0, rem = t;
  while (rem >= NT - r) {
    rem -= NT - r;
    ++r;
  }
  rt = r;
  ct = r + rem;
}

template <typename T, int KC>
__global__ __launch_bounds__(256) void k_vmat(const T* __restrict__ phi,
                                              const double* __restrict__ v, int nao_pad, int ld,
                                              int ntiles, int nwg, double* __restrict__ part) {
  typedef typename vec4_of<T>::type V4;
  extern __shared__ unsigned char smem_raw[];
  T* Ps = reinterpret_cast<T*>(smem_raw);
  T* vs = Ps + (long)nao_pad * (KC + 1);
  const int wg = blockIdx.x, grp = blockIdx.y, b = blockIdx.z;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int l16 = lane & 15, kq = lane >> 4;
  const int NT = nao_pad >> 4;
  const int tbase = grp * 4 * VMAT_TPW + w * VMAT_TPW;
  int rto[VMAT_TPW], cto[VMAT_TPW];
#pragma unroll
  for (int t = 0; t < VMAT_TPW; ++t) {
    int rt = 0, ct = 0;
    if (tbase + t < ntiles) tile_rc(tbase + t, NT, rt, ct);
    rto[t] = (rt * 16 + l16) * (KC + 1);
    cto[t] = (ct * 16 + l16) * (KC + 1);
  }
  const T* P = phi + (long)b * nao_pad * ld;
  const double* vb = v + (long)b * ld;
  f64x4 acc[VMAT_TPW];
#pragma unroll
  for (int t = 0; t < VMAT_TPW; ++t) acc[t] = f64x4{0, 0, 0, 0};
  for (long c0 = (long)wg * KC; c0 < ld; c0 += (long)nwg * KC) {
    __syncthreads();
    for (int e = threadIdx.x; e < nao_pad * KC; e += 256) {
      const int row = e / KC, kk = e - row * KC;
      Ps[row * (KC + 1) + kk] = P[(long)row * ld + c0 + kk];
    }
    for (int e = threadIdx.x; e < KC; e += 256) vs[e] = (T)vb[c0 + e];
    __syncthreads();
    V4 cacc[VMAT_TPW];
#pragma unroll
    for (int t = 0; t < VMAT_TPW; ++t) cacc[t] = V4{0, 0, 0, 0};
#pragma unroll 4
    for (int ks = 0; ks < KC / 4; ++ks) {
      const int kk = ks * 4 + kq;
      const T vk = vs[kk];
#pragma unroll
      for (int t = 0; t < VMAT_TPW; ++t) {
        if (tbase + t < ntiles) {
          const T a = Ps[rto[t] + kk];
          const T bb = vk * Ps[cto[t] + kk];
          cacc[t] = mfma16(a, bb, cacc[t]);
        }
      }
    }
#pragma unroll
    for (int t = 0; t < VMAT_TPW; ++t)
#pragma unroll
      for (int e = 0; e < 4; ++e) acc[t][e] += (double)cacc[t][e];
  }
#pragma unroll
  for (int t = 0; t < VMAT_TPW; ++t) {
    const int tile = tbase + t;
    if (tile < ntiles) {
      double* o = part + (((long)b * ntiles + tile) * nwg + wg) * 256 + lane * 4;
#pragma unroll
      for (int e = 0; e < 4; ++e) o[e] = acc[t][e];
    }
  }
}

template <typename T>
__global__ void k_vmat_reduce(const double* __restrict__ part, int nao_pad, int ntiles, int nwg,
                              double dV, double* __restrict__ V) {
  const int tile = blockIdx.x, b = blockIdx.y;
  const int idx = threadIdx.x;  // lane*4 + e
  const int lane = idx >> 2, e = idx & 3;
  const int i = acc_row<T>(lane >> 4, e), j = lane & 15;
  const double* p = part + ((long)b * ntiles + tile) * nwg * 256 + idx;
  double s = 0.0;
  for (int k = 0; k < nwg; ++k) s += p[(long)k * 256];
  s *= dV;
  int rt, ct;
  tile_rc(tile, nao_pad >> 4, rt, ct);
  const long row = rt * 16 + i, col = ct * 16 + j;
  double* Vb = V + (long)b * nao_pad * nao_pad;
  if (rt != ct) {
    Vb[row * nao_pad + col] = s;
    Vb[col * nao_pad + row] = s;
  } else if (i <= j) {
    Vb[row * nao_pad + col] = s;
    Vb[col * nao_pad + row] = s;
  }
}

static void vmat_shape(int B, int nao_pad, int ld, int prec, int& KC, int& ntiles, int& ngroups,
                       int& nwg) {
  KC = nao_pad <= 160 ? 64 : (nao_pad <= 320 ? 32 : 16);
  if (prec == 32 && nao_pad <= 320) KC = 64;
  const int NT = nao_pad >> 4;
  ntiles = NT * (NT + 1) / 2;
  ngroups = (ntiles + 4 * VMAT_TPW - 1) / (4 * VMAT_TPW);
  const int nchunks = ld / KC;
  long target = VMAT_TARGET_BLOCKS / ((long)(B > 0 ? B : 1) * ngroups);
  if (target < 1) target = 1;
  if (target > nchunks) target = nchunks;
  nwg = (int)target;
}

extern "C" int hqc_ao_vmat_workspace(int B, int nao_pad, int ld, int prec, size_t* bytes) {
  if (!bytes || B < 0 || nao_pad <= 0 || (nao_pad & 15) || ld <= 0 || (ld & 63))
    return HQC_E_ARG;
  int KC, ntiles, ngroups, nwg;
  vmat_shape(B, nao_pad, ld, prec, KC, ntiles, ngroups, nwg);
  *bytes = (size_t)B * ntiles * nwg * 256 * sizeof(double);
  return HQC_OK;
}

template <typename T>
static int launch_vmat(const T* phi, const double* v, int B, int nao_pad, int ld, double dV,
                       double* V, void* work, size_t work_bytes, int prec, void* stream) {
  if (B < 0 || nao_pad <= 0 || (nao_pad & 15) || ld <= 0 || (ld & 63)) return HQC_E_ARG;
  if (B == 0) return HQC_OK;
  int KC, ntiles, ngroups, nwg;
  vmat_shape(B, nao_pad, ld, prec, KC, ntiles, ngroups, nwg);
  const size_t need = (size_t)B * ntiles * nwg * 256 * sizeof(double);
  if (!work || work_bytes < need) return HQC_E_ARG;
  const size_t lds = ((size_t)nao_pad * (KC + 1) + KC) * sizeof(T);
  if (lds > 160 * 1024) return HQC_E_SIZE;
  hipStream_t s = (hipStream_t)stream;
  double* part = (double*)work;
  dim3 grid(nwg, ngroups, B);
  if (KC == 64)
    hipLaunchKernelGGL((k_vmat<T, 64>), grid, dim3(256), lds, s, phi, v, nao_pad, ld, ntiles,
                       nwg, part);
  else if (KC == 32)
    hipLaunchKernelGGL((k_vmat<T, 32>), grid, dim3(256), lds, s, phi, v, nao_pad, ld, ntiles,
                       nwg, part);
  else
    hipLaunchKernelGGL((k_vmat<T, 16>), grid, dim3(256), lds, s, phi, v, nao_pad, ld, ntiles,
                       nwg, part);
  HQC_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_vmat_reduce<T>, dim3(ntiles, B), dim3(256), 0, s, part, nao_pad, ntiles,
                     nwg, dV, V);
  HQC_CHECK_LAUNCH();
  return HQC_OK;
}

extern "C" int hqc_ao_vmat_f64(const double* phi, const double* v, int B, int nao_pad, int ld,
                               double dV, double* V, void* work, size_t work_bytes,
                               void* stream) {
  return launch_vmat<double>(phi, v, B, nao_pad, ld, dV, V, work, work_bytes, 64, stream);
}

extern "C" int hqc_ao_vmat_f32(const float* phi, const double* v, int B, int nao_pad, int ld,
                               double dV, double* V, void* work, size_t work_bytes,
                               void* stream) {
  return launch_vmat<float>(phi, v, B, nao_pad, ld, dV, V, work, work_bytes, 32, stream);
}
