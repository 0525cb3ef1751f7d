// AO x grid contractions: the per-iteration hot kernels (MFMA, and the vector
// ALUs for the FP64 orbital-form density).
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

#include <mutex>

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

// LDS row padding for the MFMA fragment reads.  Lane (i, kq) of a 16x16x4
// fragment reads row i, column k0 + kq; the LDS serves 32 four-byte banks per
// cycle, i.e. 16 doubles or 32 floats.  With a row stride of STR elements the
// 16 rows of one cycle fall on distinct banks when 2*STR = 2 (mod 32) for
// FP64 (STR = KC + 1) and STR = 2 (mod 32) for FP32 (STR = KC + 2, with the
// two kq values of a cycle on adjacent banks).
template <typename T>
__host__ __device__ constexpr int lds_pad() {
  return sizeof(T) == 4 ? 2 : 1;
}

// ---------------------------------------------------------------------------
// Density.  Block = 4 waves, 64 grid columns (16 per wave).  For every group of
// NTG row tiles the wave runs the full K loop over AOs: X tile = D(rows, :) *
// phi(:, cols) with D read through L2 (D is symmetric so its row-major rows
// give the A operand coalesced); the epilogue forms sum_p phi_p X_p per column.
// Serves FP32 phi; the FP64 form runs on the vector ALUs (k_density_valu).
// ---------------------------------------------------------------------------
template <typename T, int NTG>
__global__ __launch_bounds__(256) void k_density(const T* __restrict__ phi, const int* __restrict__ bidx,
                                                 const double* __restrict__ D, int nao_pad,
                                                 int ld, double* __restrict__ rho,
                                                 double* __restrict__ X) {
  typedef typename vec4_of<T>::type V4;
  const int b = blockIdx.y;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int l16 = lane & 15, kq = lane >> 4;
  const long col = (long)blockIdx.x * 64 + w * 16 + l16;
  const T* P = phi + (long)(bidx ? bidx[b] : b) * nao_pad * ld;
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
      T av[NTG];
#pragma unroll
      for (int i = 0; i < NTG; ++i) av[i] = (T)Drow[min(rt0 + i, NT - 1) * 16];
#pragma unroll
      for (int i = 0; i < NTG; ++i) acc[i] = mfma16(av[i], bv, acc[i]);
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
static int launch_density(const T* phi, const int* bidx, const double* D, int B, int nao_pad, int ld,
                          double* rho, double* X, void* stream) {
  if (B < 0 || nao_pad <= 0 || (nao_pad & 15) || ld <= 0 || (ld & 63)) return HQC_E_ARG;
  if (B == 0) return HQC_OK;
#define DENS_CASE(G)                                                                       \
  case G:                                                                                  \
    hipLaunchKernelGGL((k_density<T, G>), dim3(ld / 64, B), dim3(256), 0, (hipStream_t)stream, \
                       phi, bidx, D, nao_pad, ld, rho, X);                                 \
    break;
  switch (min(nao_pad >> 4, 8)) {
    DENS_CASE(1) DENS_CASE(2) DENS_CASE(3) DENS_CASE(4) DENS_CASE(5) DENS_CASE(6) DENS_CASE(7)
    default: DENS_CASE(8)
  }
#undef DENS_CASE
  HQC_CHECK_LAUNCH();
  return HQC_OK;
}

static int launch_density_valu(const double* phi, const int* bidx, const double* D, int B,
                               int nao_pad, int ld, double* rho, double* X, void* stream);

extern "C" int hqc_ao_density_f64(const double* phi, const int* bidx, const double* D, int B,
                                  int nao_pad, int ld, double* rho, double* X, void* stream) {
  if (B < 0 || nao_pad <= 0 || (nao_pad & 15) || ld <= 0 || (ld & 63)) return HQC_E_ARG;
  if (B == 0) return HQC_OK;
  return launch_density_valu(phi, bidx, D, B, nao_pad, ld, rho, X, stream);
}

extern "C" int hqc_ao_density_f32(const float* phi, const int* bidx, const double* D, int B,
                                  int nao_pad, int ld, double* rho, void* stream) {
  return launch_density<float>(phi, bidx, D, B, nao_pad, ld, rho, nullptr, stream);
}

// ---------------------------------------------------------------------------
// Orbital-form density rho(r) = sum_i (sum_k At[k][i] phi_k(r))^2 with
// At = C sqrt(f) restricted to the orbitals whose occupation is not
// negligible (m_pad columns, zero-padded).  Same tiling as k_density; the
// epilogue squares the accumulator instead of re-reading phi.
// ---------------------------------------------------------------------------
template <typename T, int NTG>
__global__ __launch_bounds__(256) void k_orb_density(const T* __restrict__ phi, const int* __restrict__ bidx,
                                                     const double* __restrict__ At, int nao_pad,
                                                     int m_pad, int ld,
                                                     double* __restrict__ rho) {
  typedef typename vec4_of<T>::type V4;
  const int b = blockIdx.y;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int l16 = lane & 15, kq = lane >> 4;
  const long col = (long)blockIdx.x * 64 + w * 16 + l16;
  const T* P = phi + (long)(bidx ? bidx[b] : b) * nao_pad * ld;
  const double* Ab = At + (long)b * nao_pad * m_pad;
  const int NT = m_pad >> 4, KS = nao_pad >> 2;
  double rsum = 0.0;
  for (int rt0 = 0; rt0 < NT; rt0 += NTG) {
    V4 acc[NTG];
#pragma unroll
    for (int i = 0; i < NTG; ++i) acc[i] = V4{0, 0, 0, 0};
    for (int ks = 0; ks < KS; ++ks) {
      const int k = ks * 4 + kq;
      const T bv = P[(long)k * ld + col];
      const double* Arow = Ab + (long)k * m_pad + l16;
      T av[NTG];
#pragma unroll
      for (int i = 0; i < NTG; ++i) av[i] = (T)Arow[min(rt0 + i, NT - 1) * 16];
#pragma unroll
      for (int i = 0; i < NTG; ++i) acc[i] = mfma16(av[i], bv, acc[i]);
    }
#pragma unroll
    for (int i = 0; i < NTG; ++i) {
      if (rt0 + i < NT) {
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const double y = (double)acc[i][e];
          rsum += y * y;
        }
      }
    }
  }
  rsum += __shfl_xor(rsum, 16, 64);
  rsum += __shfl_xor(rsum, 32, 64);
  if (kq == 0) rho[(long)b * ld + col] = rsum;
}

// LDS variant (coefficients of up to 4 orbital tiles staged once per workgroup).
// Workgroup = ODENS_CG groups of 256 grid columns, wave = 64 columns = 4 column
// tiles per group, so every coefficient fragment read from LDS feeds 4 MFMAs
// and every phi fragment feeds MT; phi fragments are prefetched two k-steps
// ahead.
#define ODENS_CT 4
#define ODENS_CG 4          // column groups of 256 per workgroup (one staging of At)
#define ODENS_LDS_MAX (96 * 1024)

template <typename T, int MT>
__global__ __launch_bounds__(256) void k_orb_density_lds(const T* __restrict__ phi, const int* __restrict__ bidx,
                                                         const double* __restrict__ At,
                                                         int nao_pad, int m_pad, int ld,
                                                         double* __restrict__ rho) {
  typedef typename vec4_of<T>::type V4;
  constexpr int MW = MT * 16;                   // orbitals per group
  extern __shared__ unsigned char smem_raw[];
  T* As = reinterpret_cast<T*>(smem_raw);       // [nao_pad][MW]
  const int b = blockIdx.y;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int l16 = lane & 15, kq = lane >> 4;
  const T* Pb = phi + (long)(bidx ? bidx[b] : b) * nao_pad * ld + l16;
  const double* Ab = At + (long)b * nao_pad * m_pad;
  const int KS = nao_pad >> 2;
  // per column group: the density of this lane's column (before the kq reduction)
  double rs[ODENS_CG][ODENS_CT];
#pragma unroll
  for (int g = 0; g < ODENS_CG; ++g)
#pragma unroll
    for (int c = 0; c < ODENS_CT; ++c) rs[g][c] = 0.0;
  for (int mg = 0; mg < m_pad; mg += MW) {
    __syncthreads();
    for (int e = threadIdx.x; e < nao_pad * MW; e += 256) {
      const int k = e / MW, j = e - k * MW;
      As[e] = (mg + j < m_pad) ? (T)Ab[(long)k * m_pad + mg + j] : (T)0;
    }
    __syncthreads();
#pragma unroll
    for (int g = 0; g < ODENS_CG; ++g) {
      const long col0 = ((long)blockIdx.x * ODENS_CG + g) * 256 + w * 64;
      if (col0 >= ld) continue;                 // wave-uniform (ld is a multiple of 64)
      const T* P = Pb + col0;
      V4 acc[MT][ODENS_CT];
#pragma unroll
      for (int i = 0; i < MT; ++i)
#pragma unroll
        for (int c = 0; c < ODENS_CT; ++c) acc[i][c] = V4{0, 0, 0, 0};
      T bq[2][ODENS_CT];
#pragma unroll
      for (int q = 0; q < 2; ++q)
#pragma unroll
        for (int c = 0; c < ODENS_CT; ++c)
          bq[q][c] = (q < KS) ? P[(long)(q * 4 + kq) * ld + c * 16] : (T)0;
      for (int ks = 0; ks < KS; ks += 2) {
#pragma unroll
        for (int q = 0; q < 2; ++q) {
          const int kk = ks + q;
          if (kk < KS) {
            const int k = kk * 4 + kq;
            T a[MT];
#pragma unroll
            for (int i = 0; i < MT; ++i) a[i] = As[k * MW + i * 16 + l16];
#pragma unroll
            for (int i = 0; i < MT; ++i)
#pragma unroll
              for (int c = 0; c < ODENS_CT; ++c) acc[i][c] = mfma16(a[i], bq[q][c], acc[i][c]);
            if (kk + 2 < KS) {
#pragma unroll
              for (int c = 0; c < ODENS_CT; ++c)
                bq[q][c] = P[(long)((kk + 2) * 4 + kq) * ld + c * 16];
            }
          }
        }
      }
#pragma unroll
      for (int i = 0; i < MT; ++i)
#pragma unroll
        for (int c = 0; c < ODENS_CT; ++c)
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            const double y = (double)acc[i][c][e];
            rs[g][c] += y * y;
          }
    }
  }
#pragma unroll
  for (int g = 0; g < ODENS_CG; ++g) {
    const long col0 = ((long)blockIdx.x * ODENS_CG + g) * 256 + w * 64;
    if (col0 >= ld) continue;
#pragma unroll
    for (int c = 0; c < ODENS_CT; ++c) {
      double v = rs[g][c];
      v += __shfl_xor(v, 16, 64);
      v += __shfl_xor(v, 32, 64);
      if (kq == 0) rho[(long)b * ld + col0 + c * 16 + l16] = v;
    }
  }
}

// FP64 orbital density on the vector ALUs.  On gfx950 a stream of v_fma_f64
// sustains more FP64 throughput than the 16x16x4 f64 MFMA (measured: this
// kernel 8.5 ms against 10.9 ms for the MFMA + LDS form at 54 H dzv, batch
// 256), and the left operand sqrt(f_i) C_pi is the same for every lane of a
// wave: it is read with scalar loads and enters each FMA as the scalar operand,
// so it costs neither LDS bandwidth nor vector registers.  Workgroup = NW waves
// over the same 64*VD_COLS grid columns (stride 64 between a lane's columns, so
// every load is one coalesced 512-byte row segment); wave w owns orbitals
// [mg + MG w, mg + MG (w + 1)), MG | m_pad.  The waves read the same phi rows, so
// HBM delivers phi once and L1 serves the other reads.  The per-wave partial
// sums sum_i psi_i^2 are added through LDS in a fixed order.
#define VD_COLS 2
#define VD_PF 4

template <int MG, int NW>
__global__ __launch_bounds__(64 * NW) void k_orb_density_valu(const double* __restrict__ phi,
                                                              const int* __restrict__ bidx,
                                                              const double* __restrict__ At,
                                                              int nao_pad, int m_pad, int ld,
                                                              double* __restrict__ rho) {
  constexpr int NC = 64 * VD_COLS;              // columns per workgroup
  __shared__ double red[NW][NC];
  const int b = blockIdx.y;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const long col0 = (long)blockIdx.x * NC;
  const double* P = phi + (long)(bidx ? bidx[b] : b) * nao_pad * ld + col0 + lane;
  const double* Ab = At + (long)b * nao_pad * m_pad;
  // a column block past ld (last workgroup only) re-reads block 0 instead, so
  // every load stays inside its row; those columns are never stored
  int off[VD_COLS];
#pragma unroll
  for (int j = 0; j < VD_COLS; ++j) off[j] = (col0 + 64 * j < ld) ? 64 * j : 0;
  double rs[VD_COLS];
#pragma unroll
  for (int j = 0; j < VD_COLS; ++j) rs[j] = 0.0;
  const int klast = nao_pad - 1;
  for (int mg = 0; mg < m_pad; mg += NW * MG) {
    const int o0 = mg + w * MG;
    if (o0 >= m_pad) continue;                  // wave-uniform
    const double* A = Ab + o0;
    double acc[MG][VD_COLS];
#pragma unroll
    for (int i = 0; i < MG; ++i)
#pragma unroll
      for (int j = 0; j < VD_COLS; ++j) acc[i][j] = 0.0;
    // nao_pad is a multiple of 16, so the k loop runs whole groups of VD_PF;
    // the loads past the last row re-read the last row (values unused)
    double pb[VD_PF][VD_COLS];
#pragma unroll
    for (int q = 0; q < VD_PF; ++q)
#pragma unroll
      for (int j = 0; j < VD_COLS; ++j) pb[q][j] = P[(long)q * ld + off[j]];
    for (int k0 = 0; k0 < nao_pad; k0 += VD_PF) {
#pragma unroll
      for (int q = 0; q < VD_PF; ++q) {
        const int k = k0 + q;
        const int kp = min(k + VD_PF, klast);
        double a[MG];
#pragma unroll
        for (int i = 0; i < MG; ++i) a[i] = A[(long)k * m_pad + i];
        double p[VD_COLS];
#pragma unroll
        for (int j = 0; j < VD_COLS; ++j) {
          p[j] = pb[q][j];
          pb[q][j] = P[(long)kp * ld + off[j]];
        }
#pragma unroll
        for (int i = 0; i < MG; ++i)
#pragma unroll
          for (int j = 0; j < VD_COLS; ++j) acc[i][j] = fma(a[i], p[j], acc[i][j]);
      }
    }
#pragma unroll
    for (int i = 0; i < MG; ++i)
#pragma unroll
      for (int j = 0; j < VD_COLS; ++j) rs[j] = fma(acc[i][j], acc[i][j], rs[j]);
  }
#pragma unroll
  for (int j = 0; j < VD_COLS; ++j) red[w][lane + 64 * j] = rs[j];
  __syncthreads();
  for (int t = threadIdx.x; t < NC; t += 64 * NW) {
    double s = 0.0;
#pragma unroll
    for (int v = 0; v < NW; ++v) s += red[v][t];
    if (col0 + t < ld) rho[(long)b * ld + col0 + t] = s;
  }
}

// FP64 D-form density with the AO-space product kept: X = D^T phi (D is
// symmetric, so X_p = sum_q D_pq phi_q, the force adjoint's operand) and
// rho = sum_p phi_p X_p, in the layout of k_orb_density_valu with D in place of
// At: D[k][p] is wave-uniform and enters each FMA through a scalar load, wave w
// of a pass owns the MG rows [MG (NW pass + w), +MG) of X (a last group that
// would run past nao_pad is moved back to end at nao_pad and stores only the
// rows no other group owns; MG <= nao_pad), and the waves of a
// workgroup read the same phi rows (HBM once, L1 for the rest).  The epilogue
// stores X (512-byte row segments) and adds phi_p X_p with phi re-read through
// the cache; the per-wave sums meet in LDS in a fixed order.  Against the
// 16x16x4 f64 MFMA form (k_density) at 256 geometries x 112 AOs x 49^3 points:
// 22.6 instead of 32.6 ms with X, whose every K step waited on eight L2 loads.
template <int MG, int NW>
__global__ __launch_bounds__(64 * NW) void k_density_valu(const double* __restrict__ phi,
                                                          const int* __restrict__ bidx,
                                                          const double* __restrict__ D,
                                                          int nao_pad, int ld,
                                                          double* __restrict__ rho,
                                                          double* __restrict__ X) {
  constexpr int NC = 64 * VD_COLS;
  __shared__ double red[NW][NC];
  const int b = blockIdx.y;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const long col0 = (long)blockIdx.x * NC;
  const double* P = phi + (long)(bidx ? bidx[b] : b) * nao_pad * ld + col0 + lane;
  const double* Db = D + (long)b * nao_pad * nao_pad;
  double* Xb = X ? X + (long)b * nao_pad * ld + col0 + lane : nullptr;
  // a column block past ld (last workgroup only) re-reads block 0 instead, so
  // every load stays inside its row; those columns are never stored
  int off[VD_COLS];
#pragma unroll
  for (int j = 0; j < VD_COLS; ++j) off[j] = (col0 + 64 * j < ld) ? 64 * j : 0;
  double rs[VD_COLS];
#pragma unroll
  for (int j = 0; j < VD_COLS; ++j) rs[j] = 0.0;
  const int klast = nao_pad - 1;
  for (int mg = 0; mg < nao_pad; mg += NW * MG) {
    const int o0 = mg + w * MG;
    if (o0 >= nao_pad) continue;                // wave-uniform
    const int ob = min(o0, nao_pad - MG);       // rows ob .. ob + MG - 1
    double acc[MG][VD_COLS];
#pragma unroll
    for (int i = 0; i < MG; ++i)
#pragma unroll
      for (int j = 0; j < VD_COLS; ++j) acc[i][j] = 0.0;
    double pb[VD_PF][VD_COLS];
#pragma unroll
    for (int q = 0; q < VD_PF; ++q)
#pragma unroll
      for (int j = 0; j < VD_COLS; ++j) pb[q][j] = P[(long)q * ld + off[j]];
    for (int k0 = 0; k0 < nao_pad; k0 += VD_PF) {
#pragma unroll
      for (int q = 0; q < VD_PF; ++q) {
        const int k = k0 + q;
        const int kp = min(k + VD_PF, klast);
        const double* Dk = Db + (long)k * nao_pad + ob;
        double a[MG];
#pragma unroll
        for (int i = 0; i < MG; ++i) a[i] = Dk[i];
        double p[VD_COLS];
#pragma unroll
        for (int j = 0; j < VD_COLS; ++j) {
          p[j] = pb[q][j];
          pb[q][j] = P[(long)kp * ld + off[j]];
        }
#pragma unroll
        for (int i = 0; i < MG; ++i)
#pragma unroll
          for (int j = 0; j < VD_COLS; ++j) acc[i][j] = fma(a[i], p[j], acc[i][j]);
      }
    }
#pragma unroll
    for (int i = 0; i < MG; ++i) {
      if (ob + i >= o0) {                       // wave-uniform
        const long row = (long)(ob + i) * ld;
#pragma unroll
        for (int j = 0; j < VD_COLS; ++j) {
          rs[j] = fma(acc[i][j], P[row + off[j]], rs[j]);
          if (Xb && col0 + 64 * j < ld) Xb[row + 64 * j] = acc[i][j];
        }
      }
    }
  }
#pragma unroll
  for (int j = 0; j < VD_COLS; ++j) red[w][lane + 64 * j] = rs[j];
  __syncthreads();
  for (int t = threadIdx.x; t < NC; t += 64 * NW) {
    double s = 0.0;
#pragma unroll
    for (int v = 0; v < NW; ++v) s += red[v][t];
    if (col0 + t < ld) rho[(long)b * ld + col0 + t] = s;
  }
}

static int launch_density_valu(const double* phi, const int* bidx, const double* D, int B,
                               int nao_pad, int ld, double* rho, double* X, void* stream) {
  // 32-row groups over at most 8 waves, balanced (112 AOs: 4 waves, one pass)
  const int mgr = nao_pad >= 32 ? 32 : 16;
  const int ng = (nao_pad + mgr - 1) / mgr;
  const int passes = (ng + 7) / 8;
  const int nw = (ng + passes - 1) / passes;
  const dim3 grid((ld + 64 * VD_COLS - 1) / (64 * VD_COLS), B);
  hipStream_t s = (hipStream_t)stream;
#define DV_CASE(N)                                                                         \
  case N:                                                                                  \
    hipLaunchKernelGGL((k_density_valu<32, N>), grid, dim3(64 * N), 0, s, phi, bidx, D,    \
                       nao_pad, ld, rho, X);                                               \
    break;
  if (mgr == 16) {                              // nao_pad = 16: one group
    hipLaunchKernelGGL((k_density_valu<16, 1>), grid, dim3(64), 0, s, phi, bidx, D, nao_pad,
                       ld, rho, X);
    HQC_CHECK_LAUNCH();
    return HQC_OK;
  }
  switch (nw) {
    DV_CASE(1) DV_CASE(2) DV_CASE(3) DV_CASE(4) DV_CASE(5) DV_CASE(6) DV_CASE(7)
    default: DV_CASE(8)
  }
#undef DV_CASE
  HQC_CHECK_LAUNCH();
  return HQC_OK;
}

template <typename T>
static int launch_orb_density(const T* phi, const int* bidx, const double* At, int B, int nao_pad, int m_pad,
                              int ld, double* rho, void* stream) {
  if (B < 0 || nao_pad <= 0 || (nao_pad & 15) || m_pad <= 0 || (m_pad & 15) || ld <= 0 ||
      (ld & 63))
    return HQC_E_ARG;
  if (B == 0) return HQC_OK;
  if (sizeof(T) == 8) {
    const dim3 grid((ld + 64 * VD_COLS - 1) / (64 * VD_COLS), B);
    const double* ph = reinterpret_cast<const double*>(phi);
    if (m_pad % 32 == 0)
      hipLaunchKernelGGL((k_orb_density_valu<32, 2>), grid, dim3(128), 0, (hipStream_t)stream, ph,
                         bidx, At, nao_pad, m_pad, ld, rho);
    else
      hipLaunchKernelGGL((k_orb_density_valu<16, 4>), grid, dim3(256), 0, (hipStream_t)stream, ph,
                         bidx, At, nao_pad, m_pad, ld, rho);
    HQC_CHECK_LAUNCH();
    return HQC_OK;
  }
  {
    const int mt = min(m_pad >> 4, 4);
    const size_t lds = (size_t)nao_pad * mt * 16 * sizeof(T);
    if (lds <= ODENS_LDS_MAX) {
      if (lds > 64 * 1024) {
        const void* f = mt == 4 ? (const void*)k_orb_density_lds<T, 4>
                      : mt == 3 ? (const void*)k_orb_density_lds<T, 3>
                      : mt == 2 ? (const void*)k_orb_density_lds<T, 2>
                                : (const void*)k_orb_density_lds<T, 1>;
        if (hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) !=
            hipSuccess)
          (void)hipGetLastError();
      }
      const dim3 grid((ld + 256 * ODENS_CG - 1) / (256 * ODENS_CG), B);
      switch (mt) {
        case 1:
          hipLaunchKernelGGL((k_orb_density_lds<T, 1>), grid, dim3(256), lds,
                             (hipStream_t)stream, phi, bidx, At, nao_pad, m_pad, ld, rho);
          break;
        case 2:
          hipLaunchKernelGGL((k_orb_density_lds<T, 2>), grid, dim3(256), lds,
                             (hipStream_t)stream, phi, bidx, At, nao_pad, m_pad, ld, rho);
          break;
        case 3:
          hipLaunchKernelGGL((k_orb_density_lds<T, 3>), grid, dim3(256), lds,
                             (hipStream_t)stream, phi, bidx, At, nao_pad, m_pad, ld, rho);
          break;
        default:
          hipLaunchKernelGGL((k_orb_density_lds<T, 4>), grid, dim3(256), lds,
                             (hipStream_t)stream, phi, bidx, At, nao_pad, m_pad, ld, rho);
      }
      HQC_CHECK_LAUNCH();
      return HQC_OK;
    }
  }
#define ODENS_CASE(G)                                                                      \
  case G:                                                                                  \
    hipLaunchKernelGGL((k_orb_density<T, G>), dim3(ld / 64, B), dim3(256), 0,               \
                       (hipStream_t)stream, phi, bidx, At, nao_pad, m_pad, ld, rho);             \
    break;
  switch (min(m_pad >> 4, 8)) {
    ODENS_CASE(1) ODENS_CASE(2) ODENS_CASE(3) ODENS_CASE(4) ODENS_CASE(5) ODENS_CASE(6)
    ODENS_CASE(7)
    default: ODENS_CASE(8)
  }
#undef ODENS_CASE
  HQC_CHECK_LAUNCH();
  return HQC_OK;
}

extern "C" int hqc_orb_density_f64(const double* phi, const int* bidx, const double* At, int B,
                                   int nao_pad, int m_pad, int ld, double* rho, void* stream) {
  return launch_orb_density<double>(phi, bidx, At, B, nao_pad, m_pad, ld, rho, stream);
}

extern "C" int hqc_orb_density_f32(const float* phi, const int* bidx, const double* At, int B,
                                   int nao_pad, int m_pad, int ld, double* rho, void* stream) {
  return launch_orb_density<float>(phi, bidx, At, B, nao_pad, m_pad, ld, rho, stream);
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
#define VMAT_TARGET_BLOCKS 1024
#ifndef VH_TARGET_BLOCKS
#define VH_TARGET_BLOCKS 2048   // split-product kernel: workgroups to aim for per launch
#endif
// grid chunks per geometry at most: a few geometries would otherwise be cut into
// hundreds of partial tiles each, whose reduction (one thread per element
// walking the partials) then costs more than the products (4 geometries:
// 0.37 -> 0.21 ms per V; batch 256 keeps 8 chunks per geometry)
#define VMAT_MAX_PARTS 64

__device__ __forceinline__ void tile_rc(int t, int NT, int& rt, int& ct) {
  int r = 0, rem = t;
  while (rem >= NT - r) {
    rem -= NT - r;
    ++r;
  }
  rt = r;
  ct = r + rem;
}

// Large n_ao (nao_pad > 128, e.g. 128 H dzvp or the stacked complex rows of a
// k-point): 2-D output blocking.  Workgroup = one 128x128 block pair (I <= J) of
// V; wave (wr, wc) owns a 64x64 quadrant = 4x4 MFMA tiles (the quadrant below
// the diagonal of a diagonal block is skipped).  Per grid chunk of KC points the
// 128 rows of block I and of block J are staged in LDS (double-buffered, the next
// chunk held in registers during the MFMAs), so every phi element read from HBM
// feeds 128 output columns instead of 32.
#define VBIG 128

template <typename T, int KC>
__global__ __launch_bounds__(256) void k_vmat_big(const T* __restrict__ phi, const int* __restrict__ bidx,
                                                  const double* __restrict__ v, int nao_pad,
                                                  int ld, int ntiles, int nwg,
                                                  double* __restrict__ part) {
  typedef typename vec4_of<T>::type V4;
  constexpr int VW = 16 / sizeof(T);            // elements per 16-byte vector
  constexpr int VPR = KC / VW;                  // vectors per row per chunk
  constexpr int NV = 2 * VBIG * VPR / 256;      // vectors per thread per chunk
  constexpr int STR = KC + lds_pad<T>();
  typedef T vec_t __attribute__((ext_vector_type(VW)));
  extern __shared__ unsigned char smem_raw[];
  T* Ls = reinterpret_cast<T*>(smem_raw);       // [2 buffers][2 operands][VBIG*STR]
  T* vs = Ls + 4 * VBIG * STR;                  // [2][KC]
  const int wg = blockIdx.x, bp = blockIdx.y, b = blockIdx.z;
  const int NT = nao_pad >> 4, NB = (nao_pad + VBIG - 1) / VBIG;
  int I = 0, rem = bp;
  while (rem >= NB - I) {
    rem -= NB - I;
    ++I;
  }
  const int J = I + rem;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int wr = w >> 1, wc = w & 1;
  const bool active = !(I == J && wr > wc);
  const int l16 = lane & 15, kq = lane >> 4;
  const T* P = phi + (long)(bidx ? bidx[b] : b) * nao_pad * ld;
  const double* vb = v + (long)b * ld;
  vec_t stage[NV];
  double vstage = 0.0;
  auto gload = [&](long c0) {
#pragma unroll
    for (int j = 0; j < NV; ++j) {
      const int q = threadIdx.x + 256 * j;
      const int o = q / (VBIG * VPR), row = (q / VPR) % VBIG, cv = q % VPR;
      const int grow = min((o == 0 ? I : J) * VBIG + row, nao_pad - 1);
      stage[j] = *reinterpret_cast<const vec_t*>(P + (long)grow * ld + c0 + cv * VW);
    }
    if (threadIdx.x < KC) vstage = vb[c0 + threadIdx.x];
  };
  auto lstore = [&](int buf) {
#pragma unroll
    for (int j = 0; j < NV; ++j) {
      const int q = threadIdx.x + 256 * j;
      const int o = q / (VBIG * VPR), row = (q / VPR) % VBIG, cv = q % VPR;
      T* dst = Ls + (long)(buf * 2 + o) * VBIG * STR + row * STR + cv * VW;
#pragma unroll
      for (int e = 0; e < VW; ++e) dst[e] = stage[j][e];
    }
    if (threadIdx.x < KC) vs[buf * KC + threadIdx.x] = (T)vstage;
  };
  V4 acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = V4{0, 0, 0, 0};
  long c0 = (long)wg * KC;
  const long cstep = (long)nwg * KC;
  if (c0 < ld) {
    gload(c0);
    lstore(0);
  }
  __syncthreads();
  int buf = 0;
  for (; c0 < ld; c0 += cstep) {
    const bool more = c0 + cstep < ld;
    if (more) gload(c0 + cstep);
    if (active) {
      const T* Ab = Ls + (long)(buf * 2) * VBIG * STR + (wr * 64 + l16) * STR;
      const T* Bb = Ls + (long)(buf * 2 + 1) * VBIG * STR + (wc * 64 + l16) * STR;
      const T* vsb = vs + buf * KC;
#pragma unroll 2
      for (int ks = 0; ks < KC / 4; ++ks) {
        const int kk = ks * 4 + kq;
        const T vk = vsb[kk];
        T a[4], bb[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          a[t] = Ab[t * 16 * STR + kk];
          bb[t] = vk * Bb[t * 16 * STR + kk];
        }
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
          for (int j = 0; j < 4; ++j) acc[i][j] = mfma16(a[i], bb[j], acc[i][j]);
      }
    }
    if (more) lstore(buf ^ 1);
    __syncthreads();
    buf ^= 1;
  }
  if (!active) return;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int R = I * (VBIG / 16) + wr * 4 + i, C = J * (VBIG / 16) + wc * 4 + j;
      if (R < NT && C < NT && R <= C) {
        const int tile = R * NT - (R * (R - 1)) / 2 + (C - R);
        double* o = part + (((long)b * ntiles + tile) * nwg + wg) * 256 + lane * 4;
#pragma unroll
        for (int e = 0; e < 4; ++e) o[e] = (double)acc[i][j][e];
      }
    }
  }
}

// Pipelined variant for nao_pad <= 128: two LDS buffers; while the waves run
// MFMAs on chunk i, every thread already holds its share of chunk i+1 in
// registers (16-byte global loads issued before the compute), and stores it to
// the other buffer afterwards, so one barrier per chunk separates the phases.
#define VMAT_PIPE_MAXV 8

template <typename T, int KC, int TPW>
__global__ __launch_bounds__(256) void k_vmat_pipe(const T* __restrict__ phi, const int* __restrict__ bidx,
                                                   const double* __restrict__ v, int nao_pad,
                                                   int ld, int ntiles, int nwg,
                                                   double* __restrict__ part) {
  typedef typename vec4_of<T>::type V4;
  constexpr int VW = 16 / sizeof(T);            // elements per 16-byte vector
  constexpr int VPR = KC / VW;                  // vectors per row of a chunk (16)
  constexpr int STR = KC + lds_pad<T>();
  typedef T vec_t __attribute__((ext_vector_type(VW)));
  extern __shared__ unsigned char smem_raw[];
  T* Ps = reinterpret_cast<T*>(smem_raw);                    // [2][nao_pad*STR]
  T* vs = Ps + 2 * (long)nao_pad * STR;                       // [2][KC]
  const int wg = blockIdx.x, grp = blockIdx.y, b = blockIdx.z;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int l16 = lane & 15, kq = lane >> 4;
  const int NT = nao_pad >> 4;
  const int NV = nao_pad * VPR / 256;           // vectors per thread per chunk
  const int tbase = grp * 4 * TPW + w * TPW;
  // slots past the last tile recompute tile ntiles-1 (branch-free MFMA loop);
  // their results are not stored
  int rto[TPW], cto[TPW];
#pragma unroll
  for (int t = 0; t < TPW; ++t) {
    int rt, ct;
    tile_rc(min(tbase + t, ntiles - 1), NT, rt, ct);
    rto[t] = (rt * 16 + l16) * STR;
    cto[t] = (ct * 16 + l16) * STR;
  }
  const T* P = phi + (long)(bidx ? bidx[b] : b) * nao_pad * ld;
  const double* vb = v + (long)b * ld;
  vec_t stage[VMAT_PIPE_MAXV];
  double vstage = 0.0;
  auto gload = [&](long c0) {
#pragma unroll
    for (int j = 0; j < VMAT_PIPE_MAXV; ++j) {
      if (j < NV) {
        const int q = threadIdx.x + 256 * j;
        const int row = q / VPR, cv = q - row * VPR;
        stage[j] = *reinterpret_cast<const vec_t*>(P + (long)row * ld + c0 + cv * VW);
      }
    }
    if (threadIdx.x < KC) vstage = vb[c0 + threadIdx.x];
  };
  auto lstore = [&](int buf) {
    T* dst = Ps + (long)buf * nao_pad * STR;
#pragma unroll
    for (int j = 0; j < VMAT_PIPE_MAXV; ++j) {
      if (j < NV) {
        const int q = threadIdx.x + 256 * j;
        const int row = q / VPR, cv = q - row * VPR;
#pragma unroll
        for (int e = 0; e < VW; ++e) dst[row * STR + cv * VW + e] = stage[j][e];
      }
    }
    if (threadIdx.x < KC) vs[buf * KC + threadIdx.x] = (T)vstage;
  };
  V4 acc[TPW];
#pragma unroll
  for (int t = 0; t < TPW; ++t) acc[t] = V4{0, 0, 0, 0};
  long c0 = (long)wg * KC;
  const long cstep = (long)nwg * KC;
  if (c0 < ld) {
    gload(c0);
    lstore(0);
  }
  __syncthreads();
  int buf = 0;
  for (; c0 < ld; c0 += cstep) {
    const bool more = c0 + cstep < ld;
    if (more) gload(c0 + cstep);
    const T* Pb = Ps + (long)buf * nao_pad * STR;
    const T* vsb = vs + buf * KC;
#pragma unroll 4
    for (int ks = 0; ks < KC / 4; ++ks) {
      const int kk = ks * 4 + kq;
      const T vk = vsb[kk];
      T av[TPW], bv[TPW];
#pragma unroll
      for (int t = 0; t < TPW; ++t) {
        av[t] = Pb[rto[t] + kk];
        bv[t] = Pb[cto[t] + kk];
      }
#pragma unroll
      for (int t = 0; t < TPW; ++t) acc[t] = mfma16(av[t], vk * bv[t], acc[t]);
    }
    if (more) lstore(buf ^ 1);
    __syncthreads();
    buf ^= 1;
  }
#pragma unroll
  for (int t = 0; t < TPW; ++t) {
    const int tile = tbase + t;
    if (tile < ntiles) {
      double* o = part + (((long)b * ntiles + tile) * nwg + wg) * 256 + lane * 4;
#pragma unroll
      for (int e = 0; e < 4; ++e) o[e] = (double)acc[t][e];
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

static bool vmat_pipelined(int nao_pad) { return nao_pad <= 16 * VMAT_PIPE_MAXV; }

// Tiles per wave of the pipelined kernel: with g groups of 4 waves x VMAT_TPW
// tiles needed to cover all tiles, the smallest count that covers them with g
// groups, so the groups are balanced (112 AOs: 28 tiles, 1 group of 7 per wave;
// 128 AOs: 36 tiles, 2 groups of 5 instead of 8 + 1, which the autotuner measured
// at 0.49 against 0.61 ms for 4 geometries of 61^3 with 6 per wave).
static int vmat_pipe_tpw(int nao_pad) {
  const int NT = nao_pad >> 4;
  const int nt = NT * (NT + 1) / 2;
  const int groups = (nt + 4 * VMAT_TPW - 1) / (4 * VMAT_TPW);
  const int t = (nt + 4 * groups - 1) / (4 * groups);
  return hqc_tuned(HQC_TUNE_VMAT_TPW, t < VMAT_TPW ? t : VMAT_TPW);
}

static void vmat_shape(int B, int nao_pad, int ld, int prec, int& KC, int& ntiles, int& ngroups,
                       int& nwg) {
  const int NT = nao_pad >> 4;
  ntiles = NT * (NT + 1) / 2;
  long target_blocks;
  if (vmat_pipelined(nao_pad)) {
    KC = prec == 64 ? 32 : 64;                 // 16 vectors of 16 B per row (16: 64 pairs)
    const int tpw = vmat_pipe_tpw(nao_pad);
    ngroups = (ntiles + 4 * tpw - 1) / (4 * tpw);
    target_blocks = prec == 16 ? hqc_tuned(HQC_TUNE_VH_BLOCKS, VH_TARGET_BLOCKS)
                               : hqc_tuned(HQC_TUNE_VMAT_BLOCKS, 2048);
  } else {
    KC = prec == 32 ? 32 : 16;                 // 8 vectors of 16 B per row
    const int NB = (nao_pad + VBIG - 1) / VBIG;
    ngroups = NB * (NB + 1) / 2;
    target_blocks = hqc_tuned(HQC_TUNE_VBIG_BLOCKS, VMAT_TARGET_BLOCKS);
  }
  const int nchunks = ld / KC;
  long target = target_blocks / ((long)(B > 0 ? B : 1) * ngroups);
  if (target < 1) target = 1;
  const int max_parts = hqc_tuned(HQC_TUNE_VMAT_PARTS, VMAT_MAX_PARTS);
  if (target > max_parts) target = max_parts;
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

// max_parts > 0 caps the grid chunks per geometry (each one a partial tile set
// in `work`); 0 keeps vmat_shape's choice.
template <typename T>
static int launch_vmat(const T* phi, const int* bidx, const double* v, int B, int nao_pad, int ld, double dV,
                       double* V, void* work, size_t work_bytes, int prec, void* stream,
                       int max_parts = 0) {
  if (B < 0 || nao_pad <= 0 || (nao_pad & 15) || ld <= 0 || (ld & 63)) return HQC_E_ARG;
  if (B == 0) return HQC_OK;
  int KC, ntiles, ngroups, nwg;
  vmat_shape(B, nao_pad, ld, prec, KC, ntiles, ngroups, nwg);
  if (max_parts > 0 && nwg > max_parts) nwg = max_parts;
  const size_t need = (size_t)B * ntiles * nwg * 256 * sizeof(double);
  if (!work || work_bytes < need) return HQC_E_ARG;
  hipStream_t s = (hipStream_t)stream;
  double* part = (double*)work;
  dim3 grid(nwg, ngroups, B);
  if (vmat_pipelined(nao_pad)) {
    const size_t lds = 2 * ((size_t)nao_pad * (KC + lds_pad<T>()) + KC) * sizeof(T);
    constexpr int K = sizeof(T) == 4 ? 64 : 32;
#define VMAT_PIPE_CASE(TP)                                                                  \
  case TP:                                                                                  \
    hipLaunchKernelGGL((k_vmat_pipe<T, K, TP>), grid, dim3(256), lds, s, phi, bidx, v, nao_pad, ld, \
                       ntiles, nwg, part);                                                  \
    break;
    switch (vmat_pipe_tpw(nao_pad)) {
      VMAT_PIPE_CASE(1)
      VMAT_PIPE_CASE(2)
      VMAT_PIPE_CASE(3)
      VMAT_PIPE_CASE(4)
      VMAT_PIPE_CASE(5)
      VMAT_PIPE_CASE(6)
      VMAT_PIPE_CASE(7)
      default:
      VMAT_PIPE_CASE(8)
    }
#undef VMAT_PIPE_CASE
    HQC_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_vmat_reduce<T>, dim3(ntiles, B), dim3(256), 0, s, part, nao_pad, ntiles,
                       nwg, dV, V);
    HQC_CHECK_LAUNCH();
    return HQC_OK;
  }
  constexpr int KB = sizeof(T) == 4 ? 32 : 16;
  const size_t lds = ((size_t)4 * VBIG * (KB + lds_pad<T>()) + 2 * KB) * sizeof(T);
  if (hipFuncSetAttribute((const void*)k_vmat_big<T, KB>,
                          hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) != hipSuccess)
    (void)hipGetLastError();
  hipLaunchKernelGGL((k_vmat_big<T, KB>), grid, dim3(256), lds, s, phi, bidx, v, nao_pad, ld, ntiles,
                     nwg, part);
  HQC_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_vmat_reduce<T>, dim3(ntiles, B), dim3(256), 0, s, part, nao_pad, ntiles,
                     nwg, dV, V);
  HQC_CHECK_LAUNCH();
  return HQC_OK;
}

extern "C" int hqc_ao_vmat_f64(const double* phi, const int* bidx, const double* v, int B,
                               int nao_pad, int ld, double dV, double* V, void* work,
                               size_t work_bytes, void* stream) {
  return launch_vmat<double>(phi, bidx, v, B, nao_pad, ld, dV, V, work, work_bytes, 64, stream);
}

extern "C" int hqc_ao_vmat_f32(const float* phi, const int* bidx, const double* v, int B,
                               int nao_pad, int ld, double dV, double* V, void* work,
                               size_t work_bytes, void* stream) {
  return launch_vmat<float>(phi, bidx, v, B, nao_pad, ld, dV, V, work, work_bytes, 32, stream);
}

// ---------------------------------------------------------------------------
// Allocation-free form of hqc_ao_vmat_f64, for callers that cannot allocate
// device memory (code under HIP graph capture, serving loops).  The partial
// tiles that the caller's workspace holds in hqc_ao_vmat_f64 live here in a
// file-scope device array of HQC_VMAT_SCRATCH_BYTES.  Host code never passes
// the array's name to a kernel (on the host it is only a stand-in for the
// symbol): its device address is looked up once per device with that device
// current, and that address is what the kernels receive.
//
// Every partial tile slot the reduction reads is written by the product
// kernel of the same call (an empty chunk range writes zeros), so nothing is
// cleared between calls.  Calls that overlap on one device (two streams, two
// host threads) share the array and race.
// ---------------------------------------------------------------------------
#define HQC_VMAT_SCRATCH_BYTES ((size_t)32 << 20)
#define HQC_MAX_DEVICES 64
static __device__ __align__(256) double scratch[HQC_VMAT_SCRATCH_BYTES / sizeof(double)];

static std::once_flag g_scratch_once[HQC_MAX_DEVICES];
static double* g_scratch_ptr[HQC_MAX_DEVICES];
static hipError_t g_scratch_err[HQC_MAX_DEVICES];

// Device address of `scratch` on the current device (looked up on its first use there).
static int scratch_address(double** out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess) return HQC_E_HIP;
  if (dev < 0 || dev >= HQC_MAX_DEVICES) return HQC_E_SIZE;
  std::call_once(g_scratch_once[dev], [dev] {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    g_scratch_err[dev] = e;
    g_scratch_ptr[dev] = e == hipSuccess ? static_cast<double*>(p) : nullptr;
  });
  if (g_scratch_err[dev] != hipSuccess || !g_scratch_ptr[dev]) return HQC_E_HIP;
  *out = g_scratch_ptr[dev];
  return HQC_OK;
}

extern "C" size_t hqc_ao_vmat_scratch_bytes(void) { return HQC_VMAT_SCRATCH_BYTES; }

extern "C" int hqc_ao_vmat_f64_noalloc(const double* phi, const int* bidx, const double* v, int B,
                                       int nao_pad, int ld, double dV, double* V, void* stream) {
  if (B < 0 || nao_pad <= 0 || (nao_pad & 15) || ld <= 0 || (ld & 63)) return HQC_E_ARG;
  if (B == 0) return HQC_OK;
  int KC, ntiles, ngroups, nwg;
  vmat_shape(B, nao_pad, ld, 64, KC, ntiles, ngroups, nwg);
  // one partial tile set (every upper-triangle tile of every geometry) per grid chunk
  const size_t per_part = (size_t)B * ntiles * 256 * sizeof(double);
  if (per_part > HQC_VMAT_SCRATCH_BYTES) return HQC_E_SIZE;
  const size_t cap = HQC_VMAT_SCRATCH_BYTES / per_part;
  double* part = nullptr;
  const int rc = scratch_address(&part);
  if (rc != HQC_OK) return rc;
  return launch_vmat<double>(phi, bidx, v, B, nao_pad, ld, dV, V, part, HQC_VMAT_SCRATCH_BYTES, 64,
                             stream, cap < (size_t)nwg ? (int)cap : nwg);
}

// ---------------------------------------------------------------------------
// Split-precision potential matrix: FP16 products, FP32 accumulation.
// phi is kept as FP16 pairs (hi, lo) with s phi = hi + lo to 22 significant
// bits (s = 2^k chosen by the caller so that |s phi| <= 2^14; hqc_phi_split_f16).
// While a chunk is staged in LDS, v phi is formed in FP32 from the pair and
// split the same way, with a per-geometry power-of-two scale sv (|sv v| <= 1)
// from max_r |v| (the caller's, or a reduction pass of this entry point).
// Every 16x16 tile takes hi*hi + hi*lo + lo*hi per 32 grid points on
// v_mfma_f32_16x16x32_f16 (operands: 8 consecutive k per lane, the same k order
// for both; result layout as v_mfma_f32_16x16x4f32, measured); the dropped lo*lo
// term is below 2^-22 of the product, so the error is that of FP32 accumulation.
// Chunking, tiles per wave and the deterministic reduction are k_vmat_pipe's;
// phi in HBM costs 4 bytes per value, as FP32.
// ---------------------------------------------------------------------------
typedef _Float16 h16x8 __attribute__((ext_vector_type(8)));
typedef _Float16 h16x4 __attribute__((ext_vector_type(4)));
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
#define VH_KC 64
// one LDS buffer (64.5 KB at nao_pad 112) lets two workgroups share a CU, so one
// workgroup's FP16 splitting overlaps the other's MFMAs (3.3 ms against 3.9 ms
// double-buffered at 54 H dzv, batch 256)
#define VH_NBUF 1

__device__ __forceinline__ unsigned pack_h2(_Float16 hi, _Float16 lo) {
  return (unsigned)__builtin_bit_cast(unsigned short, hi) |
         ((unsigned)__builtin_bit_cast(unsigned short, lo) << 16);
}

__global__ __launch_bounds__(256) void k_phi_split_f16(const double* __restrict__ phi, long n,
                                                       double s, unsigned* __restrict__ out) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long)gridDim.x * blockDim.x) {
    const double x = phi[i] * s;
    const _Float16 hi = (_Float16)x;
    const _Float16 lo = (_Float16)(x - (double)hi);
    out[i] = pack_h2(hi, lo);
  }
}

// vm[b] = max_r |v[b,r]|
__global__ __launch_bounds__(256) void k_vmax(const double* __restrict__ v, int ld,
                                              double* __restrict__ vm) {
  __shared__ double red[4];
  const int b = blockIdx.x;
  double m = 0.0;
  for (int r = threadIdx.x; r < ld; r += 256) m = fmax(m, fabs(v[(long)b * ld + r]));
  for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) vm[b] = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
}

// Operand preparation of the mixed-precision V step, one workgroup per
// geometry b: vmax[b] = max_r |v|, dvmax[b] = max_r |v - va|, and vv = v where
// direct[b], v - va elsewhere.  One pass over v and va in place of separate
// abs, max, subtract and select passes; the values are those of the separate
// FP64 operations (a NaN propagates into the maxima).
__device__ __forceinline__ double nanmax(double m, double a) { return (a > m || a != a) ? a : m; }

__global__ __launch_bounds__(1024) void k_mixed_v_prep(const double* __restrict__ v,
                                                       const double* __restrict__ va,
                                                       const unsigned char* __restrict__ direct,
                                                       int ld, double* __restrict__ vv,
                                                       double* __restrict__ vmax,
                                                       double* __restrict__ dvmax) {
  __shared__ double red[2][16];
  const int b = blockIdx.x;
  const bool dir = direct[b] != 0;
  const double2* v2 = reinterpret_cast<const double2*>(v + (long)b * ld);
  const double2* a2 = reinterpret_cast<const double2*>(va + (long)b * ld);
  double2* o2 = reinterpret_cast<double2*>(vv + (long)b * ld);
  double m = 0.0, dm = 0.0;
  for (int r = threadIdx.x; r < ld / 2; r += 1024) {
    const double2 x = v2[r], y = a2[r];
    const double d0 = x.x - y.x, d1 = x.y - y.y;
    m = nanmax(nanmax(m, fabs(x.x)), fabs(x.y));
    dm = nanmax(nanmax(dm, fabs(d0)), fabs(d1));
    o2[r] = dir ? x : make_double2(d0, d1);
  }
  for (int o = 32; o > 0; o >>= 1) {
    m = nanmax(m, __shfl_xor(m, o, 64));
    dm = nanmax(dm, __shfl_xor(dm, o, 64));
  }
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    red[0][w] = m;
    red[1][w] = dm;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = red[0][0], c = red[1][0];
    for (int i = 1; i < 16; ++i) {
      a = nanmax(a, red[0][i]);
      c = nanmax(c, red[1][i]);
    }
    vmax[b] = a;
    dvmax[b] = c;
  }
}

extern "C" int hqc_mixed_v_prep(const double* v, const double* va, const unsigned char* direct,
                                int B, int ld, double* vv, double* vmax, double* dvmax,
                                void* stream) {
  if (B < 0 || ld <= 0 || (ld & 1)) return HQC_E_ARG;
  if (B == 0) return HQC_OK;
  if (!v || !va || !direct || !vv || !vmax || !dvmax) return HQC_E_ARG;
  if (((uintptr_t)v | (uintptr_t)va | (uintptr_t)vv) & 15) return HQC_E_ARG;
  hipLaunchKernelGGL(k_mixed_v_prep, dim3(B), dim3(1024), 0, (hipStream_t)stream, v, va, direct,
                     ld, vv, vmax, dvmax);
  HQC_CHECK_LAUNCH();
  return HQC_OK;
}

// power-of-two scale with |s v| <= 1 (1 for v = 0)
__device__ __forceinline__ float vscale_of(double m) {
  return m > 0.0 ? (float)exp2(-ceil(log2(m))) : 1.0f;
}

template <int TPW, int NBUF>
__global__ __launch_bounds__(256) void k_vmat_h2(const unsigned* __restrict__ ph,
                                                 const int* __restrict__ bidx,
                                                 const double* __restrict__ v,
                                                 const double* __restrict__ vmax, double sphi,
                                                 int nao_pad, int ld, int ntiles, int nwg,
                                                 double* __restrict__ part) {
  constexpr int KC = VH_KC;
  constexpr int STR = KC + 8;         // halves: 144-byte rows, so the 16-byte fragment
                                      // reads of 8 consecutive rows hit distinct banks
  constexpr int VPR = KC / 4;         // 16-byte vectors (4 pairs) per row of a chunk
  extern __shared__ unsigned char smem_raw[];
  _Float16* Ls = reinterpret_cast<_Float16*>(smem_raw);  // [2 buffers][Ahi,Alo,Bhi,Blo]
  const long plane = (long)nao_pad * STR;
  const int wg = blockIdx.x, grp = blockIdx.y, b = blockIdx.z;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int l16 = lane & 15, kq = lane >> 4;
  const int NT = nao_pad >> 4;
  const int NV = nao_pad * VPR / 256;  // nao_pad / 16
  const int cv = threadIdx.x % VPR;    // the same 4 columns of every chunk for this thread
  const int tbase = grp * 4 * TPW + w * TPW;
  int rto[TPW], cto[TPW], rtile[TPW];
#pragma unroll
  for (int t = 0; t < TPW; ++t) {
    int rt, ct;
    tile_rc(min(tbase + t, ntiles - 1), NT, rt, ct);
    rtile[t] = rt;                    // wave-uniform
    rto[t] = (rt * 16 + l16) * STR + 8 * kq;
    cto[t] = (ct * 16 + l16) * STR + 8 * kq;
  }
  const unsigned* P = ph + (long)(bidx ? bidx[b] : b) * nao_pad * ld;
  const double* vb = v + (long)b * ld;
  const float svb = vscale_of(vmax[b]);
  u32x4 stage[VMAT_PIPE_MAXV];
  double vst[4];
  auto gload = [&](long c0) {
#pragma unroll
    for (int j = 0; j < VMAT_PIPE_MAXV; ++j) {
      if (j < NV) {
        const int row = (threadIdx.x + 256 * j) / VPR;
        stage[j] = *reinterpret_cast<const u32x4*>(P + (long)row * ld + c0 + cv * 4);
      }
    }
#pragma unroll
    for (int e = 0; e < 4; ++e) vst[e] = vb[c0 + cv * 4 + e];
  };
  auto lstore = [&](int buf) {
    _Float16* Ahi = Ls + (long)buf * 4 * plane;
    _Float16* Alo = Ahi + plane;
    _Float16* Bhi = Alo + plane;
    _Float16* Blo = Bhi + plane;
    float vs[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) vs[e] = (float)(vst[e] * (double)svb);
    // 1 and -1 held in registers so that hi + lo and y - yh below stay fmas,
    // which read their FP16 operands directly (v_fma_mix) and round once, as
    // the convert-and-add forms do: 5 VALU operations per value instead of 9,
    // the same bits (vmat 3.05 -> 2.93 ms at 256 geometries, 112 AOs)
    float one = 1.0f, m1 = -1.0f;
    asm volatile("" : "+v"(one), "+v"(m1));
#pragma unroll
    for (int j = 0; j < VMAT_PIPE_MAXV; ++j) {
      if (j < NV) {
        const int row = (threadIdx.x + 256 * j) / VPR;
        const int off = row * STR + cv * 4;
        h16x4 ah, al, bh, bl;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const unsigned u = stage[j][e];
          const _Float16 hi = __builtin_bit_cast(_Float16, (unsigned short)(u & 0xffffu));
          const _Float16 lo = __builtin_bit_cast(_Float16, (unsigned short)(u >> 16));
          ah[e] = hi;
          al[e] = lo;
          // y = v (hi + lo) with hi + lo rounded to FP32 first (exactly as
          // (float)hi + (float)lo), then pinned as an FP32 register value (an
          // empty asm, no memory access): left free, the compiler folds the FP16
          // rounding of hi into the multiply (v_fma_mix, one rounding of the
          // exact product) in some places and not in others, and hi + lo then
          // misses y by an FP16 ulp whenever the two roundings differ.  y - yh
          // is exact in FP32, so its fma form rounds only to FP16.
          float y = vs[e] * __builtin_fmaf((float)hi, one, (float)lo);
          asm volatile("" : "+v"(y));
          const _Float16 yh = (_Float16)y;
          bh[e] = yh;
          bl[e] = (_Float16)__builtin_fmaf((float)yh, m1, y);
        }
        *reinterpret_cast<h16x4*>(Ahi + off) = ah;
        *reinterpret_cast<h16x4*>(Alo + off) = al;
        *reinterpret_cast<h16x4*>(Bhi + off) = bh;
        *reinterpret_cast<h16x4*>(Blo + off) = bl;
      }
    }
  };
  f32x4 acc[TPW];
#pragma unroll
  for (int t = 0; t < TPW; ++t) acc[t] = f32x4{0, 0, 0, 0};
  long c0 = (long)wg * KC;
  const long cstep = (long)nwg * KC;
  if (c0 < ld) {
    gload(c0);
    lstore(0);
  }
  __syncthreads();
  int buf = 0;
  for (; c0 < ld; c0 += cstep) {
    const bool more = c0 + cstep < ld;
    if (more) gload(c0 + cstep);
    const _Float16* Ahi = Ls + (long)buf * 4 * plane;
    const _Float16* Alo = Ahi + plane;
    const _Float16* Bhi = Alo + plane;
    const _Float16* Blo = Bhi + plane;
#pragma unroll
    for (int ks = 0; ks < KC / 32; ++ks) {
      // a wave's tiles run along rows of the upper triangle, so consecutive tiles
      // mostly share their row tile: its phi fragments are read from LDS once
      h16x8 ah, al;
#pragma unroll
      for (int t = 0; t < TPW; ++t) {
        if (t == 0 || rtile[t] != rtile[t - 1]) {
          ah = *reinterpret_cast<const h16x8*>(Ahi + rto[t] + 32 * ks);
          al = *reinterpret_cast<const h16x8*>(Alo + rto[t] + 32 * ks);
        }
        const h16x8 bh = *reinterpret_cast<const h16x8*>(Bhi + cto[t] + 32 * ks);
        const h16x8 bl = *reinterpret_cast<const h16x8*>(Blo + cto[t] + 32 * ks);
        acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_f16(al, bh, acc[t], 0, 0, 0);
        acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bl, acc[t], 0, 0, 0);
        acc[t] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah, bh, acc[t], 0, 0, 0);
      }
    }
    if (NBUF == 2) {
      if (more) lstore(buf ^ 1);
      __syncthreads();
      buf ^= 1;
    } else {                          // one buffer: two blocks fit a CU
      __syncthreads();
      if (more) lstore(0);
      __syncthreads();
    }
  }
  const double inv = 1.0 / (sphi * sphi * (double)svb);
#pragma unroll
  for (int t = 0; t < TPW; ++t) {
    const int tile = tbase + t;
    if (tile < ntiles) {
      double* o = part + (((long)b * ntiles + tile) * nwg + wg) * 256 + lane * 4;
#pragma unroll
      for (int e = 0; e < 4; ++e) o[e] = (double)acc[t][e] * inv;
    }
  }
}

extern "C" int hqc_phi_split_f16(const double* phi, long n, double scale, unsigned* out,
                                 void* stream) {
  if (n < 0 || !(scale > 0.0)) return HQC_E_ARG;
  if (n == 0) return HQC_OK;
  const long blocks = hqc_div_up(n, 256);
  hipLaunchKernelGGL(k_phi_split_f16, dim3((unsigned)(blocks < 65536 ? blocks : 65536)),
                     dim3(256), 0, (hipStream_t)stream, phi, n, scale, out);
  HQC_CHECK_LAUNCH();
  return HQC_OK;
}

static size_t vmat_h2_part_bytes(int B, int nao_pad, int ld) {
  int KC, ntiles, ngroups, nwg;
  vmat_shape(B, nao_pad, ld, 16, KC, ntiles, ngroups, nwg);
  return ((size_t)B * ntiles * nwg * 256 * sizeof(double) + 255) / 256 * 256;
}

extern "C" int hqc_ao_vmat_h2_workspace(int B, int nao_pad, int ld, size_t* bytes) {
  if (!bytes || B < 0 || nao_pad <= 0 || (nao_pad & 15) || ld <= 0 || (ld & 63))
    return HQC_E_ARG;
  if (!vmat_pipelined(nao_pad)) return HQC_E_SIZE;
  *bytes = vmat_h2_part_bytes(B, nao_pad, ld) + (size_t)B * sizeof(double);
  return HQC_OK;
}

extern "C" int hqc_ao_vmat_h2(const unsigned* ph, const int* bidx, const double* v,
                              const double* vmax, int B, int nao_pad, int ld, double phi_scale,
                              double dV, double* V, void* work, size_t work_bytes, void* stream) {
  if (B < 0 || nao_pad <= 0 || (nao_pad & 15) || ld <= 0 || (ld & 63) || !(phi_scale > 0.0))
    return HQC_E_ARG;
  if (!vmat_pipelined(nao_pad)) return HQC_E_SIZE;
  if (B == 0) return HQC_OK;
  int KC, ntiles, ngroups, nwg;
  vmat_shape(B, nao_pad, ld, 16, KC, ntiles, ngroups, nwg);
  const size_t pbytes = vmat_h2_part_bytes(B, nao_pad, ld);
  if (!work || work_bytes < pbytes + (size_t)B * sizeof(double)) return HQC_E_ARG;
  hipStream_t s = (hipStream_t)stream;
  double* part = (double*)work;
  if (!vmax) {
    double* vm = (double*)((char*)work + pbytes);
    hipLaunchKernelGGL(k_vmax, dim3(B), dim3(256), 0, s, v, ld, vm);
    HQC_CHECK_LAUNCH();
    vmax = vm;
  }
  const size_t lds = (size_t)VH_NBUF * 4 * nao_pad * (VH_KC + 8) * sizeof(_Float16);
  const dim3 grid(nwg, ngroups, B);
#define VMAT_H2_CASE(TP)                                                                     \
  case TP: {                                                                                 \
    if (hipFuncSetAttribute((const void*)k_vmat_h2<TP, VH_NBUF>,                             \
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) != hipSuccess) \
      (void)hipGetLastError();                                                               \
    hipLaunchKernelGGL((k_vmat_h2<TP, VH_NBUF>), grid, dim3(256), lds, s, ph, bidx, v, vmax,  \
                       phi_scale, nao_pad, ld, ntiles, nwg, part);                           \
  } break;
  switch (vmat_pipe_tpw(nao_pad)) {
    VMAT_H2_CASE(1)
    VMAT_H2_CASE(2)
    VMAT_H2_CASE(3)
    VMAT_H2_CASE(4)
    VMAT_H2_CASE(5)
    VMAT_H2_CASE(6)
    VMAT_H2_CASE(7)
    default:
    VMAT_H2_CASE(8)
  }
#undef VMAT_H2_CASE
  HQC_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_vmat_reduce<float>, dim3(ntiles, B), dim3(256), 0, s, part, nao_pad, ntiles,
                     nwg, dV, V);
  HQC_CHECK_LAUNCH();
  return HQC_OK;
}

// ---------------------------------------------------------------------------
// Orbital-form density from the FP16 pairs of hqc_phi_split_f16:
//   rho[b,r] = sum_i (sum_k At[b,k,i] phi[b,k,r])^2
// with At split the same way (per-geometry power-of-two scale, |s At| <= 2^14),
// hi*hi + hi*lo + lo*hi on v_mfma_f32_16x16x32_f16, FP32 accumulation of psi,
// FP64 accumulation of psi^2.  Workgroup = 4 waves x 64 grid columns (4 column
// tiles per wave) over ODENS_CG column groups; the split At of up to 64
// orbitals is staged in LDS once per workgroup, orbital-major so one 16-byte
// read gives a lane its 8 consecutive AOs; the phi pairs of 8 AO rows come from
// global memory per lane (rows past nao_pad re-read the last row, whose At
// rows are zero).  Relative error of psi ~2^-22 of sum_k |At phi|.
// ---------------------------------------------------------------------------
#define OH_CT 4

__device__ __forceinline__ void split_h2(double x, _Float16& hi, _Float16& lo) {
  hi = (_Float16)x;
  lo = (_Float16)(x - (double)hi);
}

template <int MT>
__global__ __launch_bounds__(256) void k_orb_density_h2(const unsigned* __restrict__ ph,
                                                        const int* __restrict__ bidx,
                                                        const double* __restrict__ At,
                                                        int nao_pad, int m_pad, int ld,
                                                        double sphi, double* __restrict__ rho) {
  constexpr int MW = MT * 16;
  extern __shared__ unsigned char smem_raw[];
  __shared__ double red[4];
  const int KP = (nao_pad + 31) & ~31;            // AO count padded to the MFMA depth
  const int STR = KP + 8;                         // halves; 16 (mod 128) bytes per row
  _Float16* Ah = reinterpret_cast<_Float16*>(smem_raw);   // [MW][STR]
  _Float16* Al = Ah + MW * STR;
  const int b = blockIdx.y;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int l16 = lane & 15, kq = lane >> 4;
  const unsigned* Pb = ph + (long)(bidx ? bidx[b] : b) * nao_pad * ld;
  const double* Ab = At + (long)b * nao_pad * m_pad;
  // per-geometry scale of At: |sA At| <= 2^14
  double m = 0.0;
  for (int e = threadIdx.x; e < nao_pad * m_pad; e += 256) m = fmax(m, fabs(Ab[e]));
  for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
  if (lane == 0) red[w] = m;
  __syncthreads();
  m = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
  const double sA = m > 0.0 ? exp2(14.0 - ceil(log2(m))) : 1.0;
  const double inv = 1.0 / (sA * sphi * sA * sphi);
  double rs[ODENS_CG][OH_CT];
#pragma unroll
  for (int g = 0; g < ODENS_CG; ++g)
#pragma unroll
    for (int c = 0; c < OH_CT; ++c) rs[g][c] = 0.0;
  for (int mg = 0; mg < m_pad; mg += MW) {
    __syncthreads();
    for (int e = threadIdx.x; e < MW * KP; e += 256) {
      const int i = e / KP, k = e - i * KP;
      _Float16 hi = (_Float16)0, lo = (_Float16)0;
      if (k < nao_pad && mg + i < m_pad) split_h2(Ab[(long)k * m_pad + mg + i] * sA, hi, lo);
      Ah[i * STR + k] = hi;
      Al[i * STR + k] = lo;
    }
    __syncthreads();
#pragma unroll
    for (int g = 0; g < ODENS_CG; ++g) {
      const long col0 = ((long)blockIdx.x * ODENS_CG + g) * 256 + w * 64;
      if (col0 >= ld) continue;                   // wave-uniform (ld is a multiple of 64)
      const unsigned* P = Pb + col0 + l16;
      f32x4 acc[MT][OH_CT];
#pragma unroll
      for (int i = 0; i < MT; ++i)
#pragma unroll
        for (int c = 0; c < OH_CT; ++c) acc[i][c] = f32x4{0, 0, 0, 0};
      for (int k0 = 0; k0 < KP; k0 += 32) {
        unsigned u[OH_CT][8];
#pragma unroll
        for (int e = 0; e < 8; ++e) {
          const int row = min(k0 + 8 * kq + e, nao_pad - 1);
#pragma unroll
          for (int c = 0; c < OH_CT; ++c) u[c][e] = P[(long)row * ld + c * 16];
        }
        h16x8 ah[MT], al[MT];
#pragma unroll
        for (int i = 0; i < MT; ++i) {
          ah[i] = *reinterpret_cast<const h16x8*>(Ah + (i * 16 + l16) * STR + k0 + 8 * kq);
          al[i] = *reinterpret_cast<const h16x8*>(Al + (i * 16 + l16) * STR + k0 + 8 * kq);
        }
#pragma unroll
        for (int c = 0; c < OH_CT; ++c) {
          h16x8 bh, bl;
#pragma unroll
          for (int e = 0; e < 8; ++e) {
            bh[e] = __builtin_bit_cast(_Float16, (unsigned short)(u[c][e] & 0xffffu));
            bl[e] = __builtin_bit_cast(_Float16, (unsigned short)(u[c][e] >> 16));
          }
#pragma unroll
          for (int i = 0; i < MT; ++i) {
            acc[i][c] = __builtin_amdgcn_mfma_f32_16x16x32_f16(al[i], bh, acc[i][c], 0, 0, 0);
            acc[i][c] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[i], bl, acc[i][c], 0, 0, 0);
            acc[i][c] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ah[i], bh, acc[i][c], 0, 0, 0);
          }
        }
      }
#pragma unroll
      for (int i = 0; i < MT; ++i)
#pragma unroll
        for (int c = 0; c < OH_CT; ++c)
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            const double y = (double)acc[i][c][e];
            rs[g][c] += y * y;
          }
    }
  }
#pragma unroll
  for (int g = 0; g < ODENS_CG; ++g) {
    const long col0 = ((long)blockIdx.x * ODENS_CG + g) * 256 + w * 64;
    if (col0 >= ld) continue;
#pragma unroll
    for (int c = 0; c < OH_CT; ++c) {
      double v = rs[g][c];
      v += __shfl_xor(v, 16, 64);
      v += __shfl_xor(v, 32, 64);
      if (kq == 0) rho[(long)b * ld + col0 + c * 16 + l16] = v * inv;
    }
  }
}

extern "C" int hqc_orb_density_h2(const unsigned* ph, const int* bidx, const double* At, int B,
                                  int nao_pad, int m_pad, int ld, double phi_scale, double* rho,
                                  void* stream) {
  if (B < 0 || nao_pad <= 0 || (nao_pad & 15) || m_pad <= 0 || (m_pad & 15) || ld <= 0 ||
      (ld & 63) || !(phi_scale > 0.0))
    return HQC_E_ARG;
  if (B == 0) return HQC_OK;
  const int mt = min(m_pad >> 4, 4);
  const int KP = (nao_pad + 31) & ~31;
  const size_t lds = (size_t)2 * mt * 16 * (KP + 8) * sizeof(_Float16);
  if (lds > 96 * 1024) return HQC_E_SIZE;
  const dim3 grid((ld + 256 * ODENS_CG - 1) / (256 * ODENS_CG), B);
  hipStream_t s = (hipStream_t)stream;
#define OH2_CASE(M)                                                                          \
  case M: {                                                                                  \
    if (lds > 64 * 1024 &&                                                                   \
        hipFuncSetAttribute((const void*)k_orb_density_h2<M>,                                \
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) != hipSuccess) \
      (void)hipGetLastError();                                                               \
    hipLaunchKernelGGL((k_orb_density_h2<M>), grid, dim3(256), lds, s, ph, bidx, At, nao_pad, \
                       m_pad, ld, phi_scale, rho);                                           \
  } break;
  switch (mt) {
    OH2_CASE(1)
    OH2_CASE(2)
    OH2_CASE(3)
    default:
    OH2_CASE(4)
  }
#undef OH2_CASE
  HQC_CHECK_LAUNCH();
  return HQC_OK;
}
