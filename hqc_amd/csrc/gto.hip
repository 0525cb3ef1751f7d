// Periodic GTO evaluation on the real-space grid: 1-D image-summed tables,
// separable AO values, and the grid adjoint used for forces.
//
// Reference: hqc/pbc/gto.py:326-369 (tables and AO values),
// hqc/pbc/solver.py:209-219 (how the solver applies them to the grid).
#include "common.h"

#define HQC_MAX_NL 4      // powers 0..3 (l_max = 1 plus two extra for gradients)
#define HQC_MAX_NA 16     // primitives per atom
#define HQC_MAX_NLOC 16   // AOs per atom

// ---------------------------------------------------------------------------
// 1-D tables: one thread per (b, A, d, j, a).
// ---------------------------------------------------------------------------
__global__ void k_gto_tables(const double* __restrict__ xp, long total, int n, int g,
                             double L, int tmax, const double* __restrict__ alpha, int na,
                             int nl, double* __restrict__ P) {
  long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= total) return;
  const int a = (int)(idx % na);
  long t1 = idx / na;
  const int j = (int)(t1 % g);
  t1 /= g;
  const int d = (int)(t1 % 3);
  const long bA = t1 / 3;  // b*n + A
  const double x = (double)j * L / (double)g - xp[bA * 3 + d];
  const double al = alpha[a];
  double acc[HQC_MAX_NL];
#pragma unroll
  for (int l = 0; l < HQC_MAX_NL; ++l) acc[l] = 0.0;
  for (int t = -tmax; t <= tmax; ++t) {
    const double xt = x - (double)t * L;
    double p = exp(-al * xt * xt);
#pragma unroll
    for (int l = 0; l < HQC_MAX_NL; ++l) {
      if (l < nl) {
        acc[l] += p;
        p *= xt;
      }
    }
  }
  double* out = P + idx * nl;
  for (int l = 0; l < nl; ++l) out[l] = acc[l];
}

extern "C" int hqc_gto_tables(const double* xp, int B, int n, int g, double L, int tmax,
                              const double* alpha, int na, int nl, double* P, void* stream) {
  if (B < 0 || n <= 0 || g <= 0 || na <= 0 || nl <= 0 || tmax < 0) return HQC_E_ARG;
  if (nl > HQC_MAX_NL || na > HQC_MAX_NA) return HQC_E_SIZE;
  if (B == 0) return HQC_OK;
  const long total = (long)B * n * 3 * g * na;
  hipLaunchKernelGGL(k_gto_tables, dim3(hqc_div_up(total, 256)), dim3(256), 0,
                     (hipStream_t)stream, xp, total, n, g, L, tmax, alpha, na, nl, P);
  HQC_CHECK_LAUNCH();
  return HQC_OK;
}

// ---------------------------------------------------------------------------
// AO values: threads over the grid points r (AOV_PPT each) for one atom A of
// geometry b.  Coefficients and component codes are staged in LDS.
// ---------------------------------------------------------------------------
#define AOV_PPT 8

template <typename T>
__global__ __launch_bounds__(256) void k_ao_values(const double* __restrict__ P, int n, int g,
                                                   int na, int nl,
                                                   const double* __restrict__ coef,
                                                   const int* __restrict__ comp, int nloc,
                                                   int nao_pad, int ld, T* __restrict__ phi) {
  __shared__ double s_coef[HQC_MAX_NA * HQC_MAX_NLOC];
  __shared__ int s_comp[HQC_MAX_NLOC];
  for (int i = threadIdx.x; i < na * nloc; i += blockDim.x) s_coef[i] = coef[i];
  for (int i = threadIdx.x; i < nloc; i += blockDim.x) s_comp[i] = comp[i];
  __syncthreads();
  const int A = blockIdx.y, b = blockIdx.z;
  const int N = g * g * g;
  const long tab = (long)(b * n + A) * 3 * g * na * nl;
  // AOV_PPT points per thread, 256 apart: one workgroup per 256 points and atom
  // spent more on its dispatch and set-up than on its values (6.4 million
  // workgroups at 256 geometries of 54 atoms, 49^3 points: 14.4 ms)
  for (int k = 0; k < AOV_PPT; ++k) {
    const int r = (blockIdx.x * AOV_PPT + k) * blockDim.x + threadIdx.x;
    if (r >= N) break;
    const int iz = r % g, iy = (r / g) % g, ix = r / (g * g);
    const double* Px = P + tab + (long)(0 * g + ix) * na * nl;
    const double* Py = P + tab + (long)(1 * g + iy) * na * nl;
    const double* Pz = P + tab + (long)(2 * g + iz) * na * nl;
    T* out = phi + ((long)b * nao_pad + (long)A * nloc) * ld + r;
    for (int o = 0; o < nloc; ++o) {
      const int c = s_comp[o];
      const int px = (c == 0), py = (c == 1), pz = (c == 2);
      double acc = 0.0;
      for (int a = 0; a < na; ++a) {
        const double w = s_coef[a * nloc + o];
        if (w != 0.0) acc += w * Px[a * nl + px] * Py[a * nl + py] * Pz[a * nl + pz];
      }
      out[(long)o * ld] = (T)acc;
    }
  }
}

template <typename T>
static int launch_ao_values(const double* P, int B, int n, int g, int na, int nl,
                            const double* coef, const int* comp, int nloc, int nao_pad, int ld,
                            T* phi, void* stream) {
  if (B < 0 || n <= 0 || g <= 0 || na <= 0 || nl <= 0 || nloc <= 0) return HQC_E_ARG;
  if (na > HQC_MAX_NA || nloc > HQC_MAX_NLOC) return HQC_E_SIZE;
  if (nao_pad < n * nloc || ld < g * g * g) return HQC_E_ARG;
  if (B == 0) return HQC_OK;
  const int N = g * g * g;
  hipLaunchKernelGGL(k_ao_values<T>, dim3(hqc_div_up(N, 256 * AOV_PPT), n, B), dim3(256), 0,
                     (hipStream_t)stream, P, n, g, na, nl, coef, comp, nloc, nao_pad, ld, phi);
  HQC_CHECK_LAUNCH();
  return HQC_OK;
}

extern "C" int hqc_ao_values_f64(const double* P, int B, int n, int g, int na, int nl,
                                 const double* coef, const int* comp, int nloc, int nao_pad,
                                 int ld, double* phi, void* stream) {
  return launch_ao_values<double>(P, B, n, g, na, nl, coef, comp, nloc, nao_pad, ld, phi,
                                  stream);
}

extern "C" int hqc_ao_values_f32(const double* P, int B, int n, int g, int na, int nl,
                                 const double* coef, const int* comp, int nloc, int nao_pad,
                                 int ld, float* phi, void* stream) {
  return launch_ao_values<float>(P, B, n, g, na, nl, coef, comp, nloc, nao_pad, ld, phi,
                                 stream);
}

// ---------------------------------------------------------------------------
// k-point tables and AO values (hqc/pbc/gto.py:371-420): the image sum carries
// the Bloch phase, P_k[.., l] = sum_t (x - tL)^l e^{-a (x - tL)^2} e^{i k_d t L},
// so the tables and the AO values are complex.  AO values are written
// "stacked": rows [0, nao_pad) hold Re phi, rows [nao_pad, 2 nao_pad) Im phi,
// which turns every complex contraction into a real one on 2 nao_pad rows.
// ---------------------------------------------------------------------------
__global__ void k_gto_tables_k(const double* __restrict__ xp, const double* __restrict__ kpt,
                               long total, int n, int g, double L, int tmax,
                               const double* __restrict__ alpha, int na, int nl,
                               double2* __restrict__ P) {
  long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= total) return;
  const int a = (int)(idx % na);
  long t1 = idx / na;
  const int j = (int)(t1 % g);
  t1 /= g;
  const int d = (int)(t1 % 3);
  const long bA = t1 / 3;  // b*n + A
  const long b = bA / n;
  const double kd = kpt[b * 3 + d];
  const double x = (double)j * L / (double)g - xp[bA * 3 + d];
  const double al = alpha[a];
  double re[HQC_MAX_NL], im[HQC_MAX_NL];
#pragma unroll
  for (int l = 0; l < HQC_MAX_NL; ++l) re[l] = im[l] = 0.0;
  for (int t = -tmax; t <= tmax; ++t) {
    const double xt = x - (double)t * L;
    double s, c;
    sincos(kd * (double)t * L, &s, &c);
    double p = exp(-al * xt * xt);
#pragma unroll
    for (int l = 0; l < HQC_MAX_NL; ++l) {
      if (l < nl) {
        re[l] += p * c;
        im[l] += p * s;
        p *= xt;
      }
    }
  }
  double2* out = P + idx * nl;
  for (int l = 0; l < nl; ++l) out[l] = make_double2(re[l], im[l]);
}

extern "C" int hqc_gto_tables_k(const double* xp, const double* kpt, int B, int n, int g,
                                double L, int tmax, const double* alpha, int na, int nl,
                                double* P, void* stream) {
  if (B < 0 || n <= 0 || g <= 0 || na <= 0 || nl <= 0 || tmax < 0 || !kpt) return HQC_E_ARG;
  if (nl > HQC_MAX_NL || na > HQC_MAX_NA) return HQC_E_SIZE;
  if (B == 0) return HQC_OK;
  const long total = (long)B * n * 3 * g * na;
  hipLaunchKernelGGL(k_gto_tables_k, dim3(hqc_div_up(total, 256)), dim3(256), 0,
                     (hipStream_t)stream, xp, kpt, total, n, g, L, tmax, alpha, na, nl,
                     (double2*)P);
  HQC_CHECK_LAUNCH();
  return HQC_OK;
}

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

__global__ __launch_bounds__(256) void k_ao_values_k(const double2* __restrict__ P, int n,
                                                     int g, int na, int nl,
                                                     const double* __restrict__ coef,
                                                     const int* __restrict__ comp, int nloc,
                                                     int nao_pad, int ld,
                                                     double* __restrict__ phi) {
  __shared__ double s_coef[HQC_MAX_NA * HQC_MAX_NLOC];
  __shared__ int s_comp[HQC_MAX_NLOC];
  for (int i = threadIdx.x; i < na * nloc; i += blockDim.x) s_coef[i] = coef[i];
  for (int i = threadIdx.x; i < nloc; i += blockDim.x) s_comp[i] = comp[i];
  __syncthreads();
  const int A = blockIdx.y, b = blockIdx.z;
  const int N = g * g * g;
  const long tab = (long)(b * n + A) * 3 * g * na * nl;
  for (int k = 0; k < AOV_PPT; ++k) {           // AOV_PPT points per thread, as k_ao_values
    const int r = (blockIdx.x * AOV_PPT + k) * blockDim.x + threadIdx.x;
    if (r >= N) break;
    const int iz = r % g, iy = (r / g) % g, ix = r / (g * g);
    const double2* Px = P + tab + (long)(0 * g + ix) * na * nl;
    const double2* Py = P + tab + (long)(1 * g + iy) * na * nl;
    const double2* Pz = P + tab + (long)(2 * g + iz) * na * nl;
    double* out_re = phi + ((long)b * 2 * nao_pad + (long)A * nloc) * ld + r;
    double* out_im = out_re + (long)nao_pad * ld;
    for (int o = 0; o < nloc; ++o) {
      const int c = s_comp[o];
      const int px = (c == 0), py = (c == 1), pz = (c == 2);
      double2 acc = make_double2(0.0, 0.0);
      for (int a = 0; a < na; ++a) {
        const double w = s_coef[a * nloc + o];
        if (w != 0.0) {
          const double2 t = cmul(cmul(Px[a * nl + px], Py[a * nl + py]), Pz[a * nl + pz]);
          acc.x += w * t.x;
          acc.y += w * t.y;
        }
      }
      out_re[(long)o * ld] = acc.x;
      out_im[(long)o * ld] = acc.y;
    }
  }
}

extern "C" int hqc_ao_values_k(const double* P, int B, int n, int g, int na, int nl,
                               const double* coef, const int* comp, int nloc, int nao_pad,
                               int ld, double* phi, void* stream) {
  if (B < 0 || n <= 0 || g <= 0 || na <= 0 || nl <= 0 || nloc <= 0) return HQC_E_ARG;
  if (na > HQC_MAX_NA || nloc > HQC_MAX_NLOC) return HQC_E_SIZE;
  if (nao_pad < n * nloc || ld < g * g * g) return HQC_E_ARG;
  if (B == 0) return HQC_OK;
  const int N = g * g * g;
  hipLaunchKernelGGL(k_ao_values_k, dim3(hqc_div_up(N, 256 * AOV_PPT), n, B), dim3(256), 0,
                     (hipStream_t)stream, (const double2*)P, n, g, na, nl, coef, comp, nloc,
                     nao_pad, ld, phi);
  HQC_CHECK_LAUNCH();
  return HQC_OK;
}

// ---------------------------------------------------------------------------
// AO values at arbitrary points (electron positions), make_pbc_gto's
// eval_pbc_ao / eval_pbc_ao_kpt (hqc/pbc/gto.py:341-369, 388-420): for r =
// x_e - R_A the 1-D image sums over t = -tmax..tmax, times the Bloch phase
// e^{i k_d t L} at a k-point.  One thread per (b, e, A); out (B, ne, nao)
// real, or complex interleaved when kpt != NULL.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_ao_points(
    const double* __restrict__ xp, const double* __restrict__ xe, const double* __restrict__ kpt,
    long total, int n, int ne, double L, int tmax, const double* __restrict__ alpha, int na,
    int nl, const double* __restrict__ coef, const int* __restrict__ comp, int nloc,
    double* __restrict__ out) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= total) return;
  const int A = (int)(idx % n);
  const long be = idx / n;             // b * ne + e
  const long b = be / ne;
  const double* x = xe + be * 3;
  const double* R = xp + (b * n + A) * 3;
  const bool cplx = kpt != nullptr;
  // 1-D tables: P[d][a][l] (re, im)
  double Pr[3][HQC_MAX_NA][2], Pi[3][HQC_MAX_NA][2];
  for (int d = 0; d < 3; ++d) {
    const double r = x[d] - R[d];
    const double kd = cplx ? kpt[b * 3 + d] : 0.0;
    for (int a = 0; a < na; ++a) {
      double sr0 = 0.0, sr1 = 0.0, si0 = 0.0, si1 = 0.0;
      for (int t = -tmax; t <= tmax; ++t) {
        const double xt = r - (double)t * L;
        const double e = exp(-alpha[a] * xt * xt);
        double c = 1.0, s = 0.0;
        if (cplx) sincos(kd * (double)t * L, &s, &c);
        sr0 += e * c;
        si0 += e * s;
        sr1 += e * xt * c;
        si1 += e * xt * s;
      }
      Pr[d][a][0] = sr0;
      Pr[d][a][1] = sr1;
      Pi[d][a][0] = si0;
      Pi[d][a][1] = si1;
    }
  }
  const long nao = (long)n * nloc;
  for (int o = 0; o < nloc; ++o) {
    const int c = comp[o];
    double ar = 0.0, ai = 0.0;
    for (int a = 0; a < na; ++a) {
      const double w = coef[a * nloc + o];
      if (w == 0.0) continue;
      double vr = 1.0, vi = 0.0;
      for (int d = 0; d < 3; ++d) {
        const int pw = (c == d) ? 1 : 0;
        const double qr = Pr[d][a][pw], qi = Pi[d][a][pw];
        const double tr = vr * qr - vi * qi;
        vi = vr * qi + vi * qr;
        vr = tr;
      }
      ar += w * vr;
      ai += w * vi;
    }
    const long oi = be * nao + (long)A * nloc + o;
    if (cplx) {
      out[2 * oi] = ar;
      out[2 * oi + 1] = ai;
    } else {
      out[oi] = ar;
    }
  }
}

extern "C" int hqc_ao_points(const double* xp, const double* xe, const double* kpt, int B, int n,
                             int ne, double L, int tmax, const double* alpha, int na, int nl,
                             const double* coef, const int* comp, int nloc, double* out,
                             void* stream) {
  if (B < 0 || n <= 0 || ne < 0 || na <= 0 || nloc <= 0 || tmax < 0 || !(L > 0.0))
    return HQC_E_ARG;
  if (na > HQC_MAX_NA || nloc > HQC_MAX_NLOC || nl > 2) return HQC_E_SIZE;
  if (B == 0 || ne == 0) return HQC_OK;
  const long total = (long)B * ne * n;
  hipLaunchKernelGGL(k_ao_points, dim3(hqc_div_up(total, 256)), dim3(256), 0,
                     (hipStream_t)stream, xp, xe, kpt, total, n, ne, L, tmax, alpha, na, nl, coef,
                     comp, nloc, out);
  HQC_CHECK_LAUNCH();
  return HQC_OK;
}

// ---------------------------------------------------------------------------
// Grid adjoint for forces.  For atom A of geometry b, each thread takes grid
// points r (grid-stride), recomputes grad phi_p(r) for the AOs of A from the
// tables (d/dx [x^l e^{-a x^2}] = l x^{l-1} e - 2a x^{l+1} e, image-summed),
// and accumulates Z_p(r) grad phi_p(r) with Z = 2 v X (+ Zx).  Each block
// writes its three partial sums; the host adds them in a fixed order.
// ---------------------------------------------------------------------------
#define HQC_FORCE_BLOCKS 16

__global__ __launch_bounds__(256) void k_grid_forces(
    const double* __restrict__ P, int n, int g, int na, int nl,
    const double* __restrict__ alpha, const double* __restrict__ coef,
    const int* __restrict__ comp, int nloc, int nao_pad, int ld,
    const double* __restrict__ X, const double* __restrict__ v,
    const double* __restrict__ Zx, double dV, int tab_lds, double* __restrict__ part) {
  __shared__ double s_coef[HQC_MAX_NA * HQC_MAX_NLOC];
  __shared__ double s_alpha[HQC_MAX_NA];
  __shared__ int s_comp[HQC_MAX_NLOC];
  __shared__ double red[4];
  extern __shared__ double s_tab[];   // this atom's 1-D tables when they fit (tab_lds)
  for (int i = threadIdx.x; i < na * nloc; i += blockDim.x) s_coef[i] = coef[i];
  for (int i = threadIdx.x; i < na; i += blockDim.x) s_alpha[i] = alpha[i];
  for (int i = threadIdx.x; i < nloc; i += blockDim.x) s_comp[i] = comp[i];
  const int A = blockIdx.y, b = blockIdx.z;
  const int N = g * g * g;
  const long tab = (long)(b * n + A) * 3 * g * na * nl;
  // the workgroup's atom tables (3 g na nl doubles, 14 KB at 49^3 dzv) staged in LDS:
  // read through L1 they thrash when the workgroups on a CU serve different atoms
  const double* T = P + tab;
  if (tab_lds) {
    for (int i = threadIdx.x; i < 3 * g * na * nl; i += blockDim.x) s_tab[i] = P[tab + i];
    T = s_tab;
  }
  __syncthreads();
  double f[3] = {0.0, 0.0, 0.0};
  for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < N; r += gridDim.x * blockDim.x) {
    const int iz = r % g, iy = (r / g) % g, ix = r / (g * g);
    const double* Pd[3] = {T + (long)(0 * g + ix) * na * nl,
                           T + (long)(1 * g + iy) * na * nl,
                           T + (long)(2 * g + iz) * na * nl};
    const double vr = v ? 2.0 * v[(long)b * ld + r] : 0.0;
    for (int o = 0; o < nloc; ++o) {
      const long row = (long)b * nao_pad + (long)A * nloc + o;
      double z = 0.0;
      if (v) z = vr * X[row * ld + r];
      if (Zx) z += Zx[row * ld + r];
      const int c = s_comp[o];
      double gr[3] = {0.0, 0.0, 0.0};
      for (int a = 0; a < na; ++a) {
        const double w = s_coef[a * nloc + o];
        if (w == 0.0) continue;
        const double m2a = -2.0 * s_alpha[a];
        double val[3], der[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          const int pw = (c == d) ? 1 : 0;
          const double* t = Pd[d] + a * nl;
          val[d] = t[pw];
          der[d] = m2a * t[pw + 1] + (pw ? t[0] : 0.0);
        }
        gr[0] += w * der[0] * val[1] * val[2];
        gr[1] += w * val[0] * der[1] * val[2];
        gr[2] += w * val[0] * val[1] * der[2];
      }
      f[0] += z * gr[0];
      f[1] += z * gr[1];
      f[2] += z * gr[2];
    }
  }
  for (int c = 0; c < 3; ++c) {
    const double t = block_sum(f[c], red);
    if (threadIdx.x == 0)
      part[(((long)b * n + A) * gridDim.x + blockIdx.x) * 3 + c] = t * dV;
  }
}

extern "C" int hqc_force_blocks(int g) {
  (void)g;
  return HQC_FORCE_BLOCKS;
}

extern "C" int hqc_grid_forces(const double* P, int B, int n, int g, int na, int nl,
                               const double* alpha, const double* coef, const int* comp,
                               int nloc, int nao_pad, int ld, const double* X, const double* v,
                               const double* Zx, double dV, double* part, void* stream) {
  if (B < 0 || n <= 0 || g <= 0 || na <= 0 || nloc <= 0) return HQC_E_ARG;
  if (na > HQC_MAX_NA || nloc > HQC_MAX_NLOC) return HQC_E_SIZE;
  if (nl < 3) return HQC_E_ARG;   // gradients need l_max + 2 powers (>= 2 for s)
  if (v && !X) return HQC_E_ARG;
  if (B == 0) return HQC_OK;
  const size_t tab_bytes = (size_t)3 * g * na * nl * sizeof(double);
  const int tab_lds = tab_bytes <= 48 * 1024;
  hipLaunchKernelGGL(k_grid_forces, dim3(HQC_FORCE_BLOCKS, n, B), dim3(256),
                     tab_lds ? tab_bytes : 0, (hipStream_t)stream, P, n, g, na, nl, alpha, coef,
                     comp, nloc, nao_pad, ld, X, v, Zx, dV, tab_lds, part);
  HQC_CHECK_LAUNCH();
  return HQC_OK;
}

// ---------------------------------------------------------------------------
// Grid adjoint at a k-point.  With stacked rows (Re phi, Im phi) and the
// stacked X = M phi of the density kernel (M = [[Re D, Im D], [-Im D, Re D]]),
// rho = sum_rows phi X and the explicit derivative of E with respect to R_A is
// the sum over both halves of Z grad(row), Z = 2 v X (+ Zx): each thread
// forms the complex gradient of phi_p from the complex tables and adds
// Z_re . grad Re phi_p + Z_im . grad Im phi_p.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_grid_forces_k(
    const double2* __restrict__ P, int n, int g, int na, int nl,
    const double* __restrict__ alpha, const double* __restrict__ coef,
    const int* __restrict__ comp, int nloc, int nao_pad, int ld,
    const double* __restrict__ X, const double* __restrict__ v,
    const double* __restrict__ Zx, double dV, double* __restrict__ part) {
  __shared__ double s_coef[HQC_MAX_NA * HQC_MAX_NLOC];
  __shared__ double s_alpha[HQC_MAX_NA];
  __shared__ int s_comp[HQC_MAX_NLOC];
  __shared__ double red[4];
  for (int i = threadIdx.x; i < na * nloc; i += blockDim.x) s_coef[i] = coef[i];
  for (int i = threadIdx.x; i < na; i += blockDim.x) s_alpha[i] = alpha[i];
  for (int i = threadIdx.x; i < nloc; i += blockDim.x) s_comp[i] = comp[i];
  __syncthreads();
  const int A = blockIdx.y, b = blockIdx.z;
  const int N = g * g * g;
  const long tab = (long)(b * n + A) * 3 * g * na * nl;
  double f[3] = {0.0, 0.0, 0.0};
  for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < N; r += gridDim.x * blockDim.x) {
    const int iz = r % g, iy = (r / g) % g, ix = r / (g * g);
    const double2* Pd[3] = {P + tab + (long)(0 * g + ix) * na * nl,
                            P + tab + (long)(1 * g + iy) * na * nl,
                            P + tab + (long)(2 * g + iz) * na * nl};
    const double vr = v ? 2.0 * v[(long)b * ld + r] : 0.0;
    for (int o = 0; o < nloc; ++o) {
      const long row = (long)b * 2 * nao_pad + (long)A * nloc + o;
      const long rowi = row + nao_pad;
      double zr = 0.0, zi = 0.0;
      if (v) {
        zr = vr * X[row * ld + r];
        zi = vr * X[rowi * ld + r];
      }
      if (Zx) {
        zr += Zx[row * ld + r];
        zi += Zx[rowi * ld + r];
      }
      const int c = s_comp[o];
      double2 gr[3] = {make_double2(0.0, 0.0), make_double2(0.0, 0.0), make_double2(0.0, 0.0)};
      for (int a = 0; a < na; ++a) {
        const double w = s_coef[a * nloc + o];
        if (w == 0.0) continue;
        const double m2a = -2.0 * s_alpha[a];
        double2 val[3], der[3];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          const int pw = (c == d) ? 1 : 0;
          const double2* t = Pd[d] + a * nl;
          val[d] = t[pw];
          der[d] = make_double2(m2a * t[pw + 1].x + (pw ? t[0].x : 0.0),
                                m2a * t[pw + 1].y + (pw ? t[0].y : 0.0));
        }
        const double2 g0 = cmul(cmul(der[0], val[1]), val[2]);
        const double2 g1 = cmul(cmul(val[0], der[1]), val[2]);
        const double2 g2 = cmul(cmul(val[0], val[1]), der[2]);
        gr[0].x += w * g0.x; gr[0].y += w * g0.y;
        gr[1].x += w * g1.x; gr[1].y += w * g1.y;
        gr[2].x += w * g2.x; gr[2].y += w * g2.y;
      }
#pragma unroll
      for (int d = 0; d < 3; ++d) f[d] += zr * gr[d].x + zi * gr[d].y;
    }
  }
  for (int c = 0; c < 3; ++c) {
    const double t = block_sum(f[c], red);
    if (threadIdx.x == 0)
      part[(((long)b * n + A) * gridDim.x + blockIdx.x) * 3 + c] = t * dV;
  }
}

extern "C" int hqc_grid_forces_k(const double* P, int B, int n, int g, int na, int nl,
                                 const double* alpha, const double* coef, const int* comp,
                                 int nloc, int nao_pad, int ld, const double* X, const double* v,
                                 const double* Zx, double dV, double* part, void* stream) {
  if (B < 0 || n <= 0 || g <= 0 || na <= 0 || nloc <= 0) return HQC_E_ARG;
  if (na > HQC_MAX_NA || nloc > HQC_MAX_NLOC) return HQC_E_SIZE;
  if (nl < 3) return HQC_E_ARG;
  if (v && !X) return HQC_E_ARG;
  if (B == 0) return HQC_OK;
  hipLaunchKernelGGL(k_grid_forces_k, dim3(HQC_FORCE_BLOCKS, n, B), dim3(256), 0,
                     (hipStream_t)stream, (const double2*)P, n, g, na, nl, alpha, coef, comp,
                     nloc, nao_pad, ld, X, v, Zx, dV, part);
  HQC_CHECK_LAUNCH();
  return HQC_OK;
}
