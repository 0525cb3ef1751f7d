// Pair densities for the Hartree-Fock exchange matrix.
//
// Reference: K_pq = sum_rs D_rs (eris + eris0)[p,q,s,r] with
// eris[p,r,s,q] = sum_G VG rhoG_rs rhoG_qp (hqc/pbc/solver.py:430-431,
// 604-614).  With D = sum_i f_i psi_i psi_i^T and
// A_ip(G) = sqrt(f_i) dV sum_r phi_p(r) psi_i(r) e^{-iG.r}, the G != 0 part is
//   K_pq = sum_i sum_G VG Re[A_ip(G) conj(A_iq(G))].
// phi_p psi_i is real, so A(-G) = conj(A(G)) and VG(-G) = VG(G): the sum over
// the symmetric odd G-set is twice the sum over one half H of it (G = 0 has
// VG = 0).  Two real pair densities share one complex FFT,
//   Z = FFT(a + i b),  A_a(G) = (Z(G) + conj Z(-G)) / 2,
//                      A_b(G) = (Z(G) - conj Z(-G)) / (2i),
// and each orbital contributes the real row block M_i = [Re A_i(H) | Im A_i(H)]
// of width 2|H| = N - 1, so K = sum_i M_i diag(2VG_H, 2VG_H) M_i^T, the
// AO x grid contraction of the potential matrix.
//
// H is the first half of the grid in centred order: column c < (N-1)/2 is the
// point with signed frequency s = (c / g^2, (c / g) % g, c % g) - g/2; its
// partner -s is column N-1-c, and the centre (N-1)/2 is G = 0.
#include <algorithm>

#include "common.h"

// z[m, p, r] = phi_p(r) (psi_{i_m}(r) + i psi_{j_m}(r)); psi rows are already
// weighted by sqrt(f) dV.  table[m] = (geometry, i, j), j < 0 for none.
// Z is double2, or float2 for the FP32 pair transforms (one rounding of the FP64
// product).
__device__ __forceinline__ void store_c(double2* z, double a, double b) { *z = make_double2(a, b); }
__device__ __forceinline__ void store_c(float2* z, double a, double b) {
  *z = make_float2((float)a, (float)b);
}

template <typename Z>
__global__ __launch_bounds__(256) void k_pair_pack(const double* __restrict__ phi,
                                                   const double* __restrict__ psi,
                                                   const int* __restrict__ table, int nao,
                                                   int nao_pad, int nmo, int N, long ld,
                                                   Z* __restrict__ z) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= N) return;
  const int p = blockIdx.y, m = blockIdx.z;
  const int b = table[3 * m], i = table[3 * m + 1], j = table[3 * m + 2];
  const double f = phi[((long)b * nao_pad + p) * ld + r];
  const double* ps = psi + (long)b * nmo * ld;
  const double a = ps[(long)i * ld + r];
  const double c = j >= 0 ? ps[(long)j * ld + r] : 0.0;
  store_c(z + ((long)m * nao + p) * N + r, f * a, f * c);
}

// One thread per column pair (c, c + nh): Z(G) and Z(-G) are read once for the
// four outputs (Re and Im of A_a and A_b); columns past 2 nh and rows past n_ao
// are written as zero.
template <typename T, typename Z>
__global__ __launch_bounds__(256) void k_pair_split(const Z* __restrict__ z, int nao,
                                                    int nao_pad, int g, int ld,
                                                    T* __restrict__ M) {
  const int N = g * g * g, nh = (N - 1) / 2;
  const int col = blockIdx.x * blockDim.x + threadIdx.x;
  const int p = blockIdx.y, m = blockIdx.z;
  if (col >= max(nh, ld - 2 * nh)) return;        // threads per row
  T* Ma = M + ((long)(2 * m) * nao_pad + p) * ld;
  T* Mb = M + ((long)(2 * m + 1) * nao_pad + p) * ld;
  // zero tail: columns [2 nh, ld)
  if (2 * nh + col < ld) {
    Ma[2 * nh + col] = (T)0;
    Mb[2 * nh + col] = (T)0;
  }
  if (col >= nh) return;
  if (p >= nao) {
    Ma[col] = Ma[col + nh] = (T)0;
    Mb[col] = Mb[col + nh] = (T)0;
    return;
  }
  const int h = g / 2;
  const int s1 = col / (g * g) - h, s2 = (col / g) % g - h, s3 = col % g - h;
  const int k = ((s1 + g) % g * g + (s2 + g) % g) * g + (s3 + g) % g;
  const int kn = ((g - s1) % g * g + (g - s2) % g) * g + (g - s3) % g;
  const Z zp = z[((long)m * nao + p) * N + k];
  const Z zn = z[((long)m * nao + p) * N + kn];
  Ma[col] = (T)(0.5 * ((double)zp.x + (double)zn.x));         // Re A_a
  Mb[col] = (T)(0.5 * ((double)zp.y + (double)zn.y));         // Re A_b
  Ma[col + nh] = (T)(0.5 * ((double)zp.y - (double)zn.y));    // Im A_a
  Mb[col + nh] = (T)(0.5 * ((double)zn.x - (double)zp.x));    // Im A_b
}

template <typename Z>
static int launch_pack(const double* phi, const double* psi, const int* table, int mp, int nao,
                       int nao_pad, int nmo, int g, int ld, Z* z, void* stream) {
  if (mp < 0 || nao <= 0 || nao > nao_pad || nmo <= 0 || g <= 0) return HQC_E_ARG;
  const long N = (long)g * g * g;
  if (ld < N || mp > 65535 || nao > 65535) return HQC_E_SIZE;
  if (mp == 0) return HQC_OK;
  hipLaunchKernelGGL(k_pair_pack<Z>, dim3(hqc_div_up(N, 256), nao, mp), dim3(256), 0,
                     (hipStream_t)stream, phi, psi, table, nao, nao_pad, nmo, (int)N, ld, z);
  HQC_CHECK_LAUNCH();
  return HQC_OK;
}

extern "C" int hqc_pair_pack(const double* phi, const double* psi, const int* table, int mp,
                             int nao, int nao_pad, int nmo, int g, int ld, double* z,
                             void* stream) {
  return launch_pack<double2>(phi, psi, table, mp, nao, nao_pad, nmo, g, ld, (double2*)z, stream);
}

extern "C" int hqc_pair_pack_c64(const double* phi, const double* psi, const int* table, int mp,
                                 int nao, int nao_pad, int nmo, int g, int ld, float* z,
                                 void* stream) {
  return launch_pack<float2>(phi, psi, table, mp, nao, nao_pad, nmo, g, ld, (float2*)z, stream);
}

template <typename T, typename Z>
static int launch_split(const Z* z, int mp, int nao, int nao_pad, int g, int ld, T* M,
                        void* stream) {
  if (mp < 0 || nao <= 0 || nao > nao_pad || g <= 0 || (g & 1) == 0) return HQC_E_ARG;
  const long N = (long)g * g * g;
  if (ld < N || mp > 65535 || nao_pad > 65535) return HQC_E_SIZE;
  if (mp == 0) return HQC_OK;
  const long nthr = std::max((N - 1) / 2, ld - (N - 1));     // threads per row
  hipLaunchKernelGGL((k_pair_split<T, Z>), dim3(hqc_div_up(nthr, 256), nao_pad, mp), dim3(256),
                     0, (hipStream_t)stream, z, nao, nao_pad, g, ld, M);
  HQC_CHECK_LAUNCH();
  return HQC_OK;
}

extern "C" int hqc_pair_split_f64(const double* z, int mp, int nao, int nao_pad, int g, int ld,
                                  double* M, void* stream) {
  return launch_split<double>((const double2*)z, mp, nao, nao_pad, g, ld, M, stream);
}

extern "C" int hqc_pair_split_f32(const double* z, int mp, int nao, int nao_pad, int g, int ld,
                                  float* M, void* stream) {
  return launch_split<float>((const double2*)z, mp, nao, nao_pad, g, ld, M, stream);
}

extern "C" int hqc_pair_split_c64(const float* z, int mp, int nao, int nao_pad, int g, int ld,
                                  float* M, void* stream) {
  return launch_split<float>((const float2*)z, mp, nao, nao_pad, g, ld, M, stream);
}
