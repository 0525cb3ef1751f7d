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
                                                   int nao_pad, int ld, T* __restrict__ phi,
                                                   unsigned long long* __restrict__ amax) {
  __shared__ double s_coef[HQC_MAX_NA * HQC_MAX_NLOC];
  __shared__ int s_comp[HQC_MAX_NLOC];
  __shared__ double red[4];
  double mx = 0.0;
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
      mx = fmax(mx, fabs(acc));
    }
  }
  if (amax) {
    const double m = block_max(mx, red);
    if (threadIdx.x == 0) atomicMax(amax + b, (unsigned long long)__double_as_longlong(m));
  }
}

// AO values along grid lines: lane = iz, so each store is one grid line of an AO
// (a coalesced row segment), each lane keeps its z factors t_z(iz)[a][0..1] in
// registers for the whole launch, and a line's x and y factors are the same for
// every lane (scalar loads).  The pointwise form above reads its z factors per
// point with lanes 8*nl bytes apart, many cache lines per wave load; this form
// reads them once.  Workgroup = AOV_LNW waves on the lines [AOV_LINES*x,
// +AOV_LINES) of one atom; ZC = ceil(g / 64) lane chunks; na <= AOV_LMAX_NA.
#define AOV_LNW 4
#define AOV_LINES 256
#define AOV_LMAX_NA 8

template <typename T, int ZC>
__global__ __launch_bounds__(64 * AOV_LNW) void k_ao_values_lines(
    const double* __restrict__ P, int n, int g, int na, int nl,
    const double* __restrict__ coef, const int* __restrict__ comp, int nloc, int nao_pad,
    int ld, T* __restrict__ phi, unsigned long long* __restrict__ amax) {
  __shared__ double red[AOV_LNW];
  const int A = blockIdx.y, b = blockIdx.z;
  const int lane = threadIdx.x & 63;
  double mx = 0.0;                                   // max |phi| of this thread's values
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const long tstride = (long)g * na * nl;
  const double* Tx = P + (long)(b * n + A) * 3 * tstride;
  const double* Ty = Tx + tstride;
  const double* Tz = Ty + tstride;
  double tz[ZC][AOV_LMAX_NA][2];
#pragma unroll
  for (int zc = 0; zc < ZC; ++zc) {
    const int iz = min(lane + 64 * zc, g - 1);    // lanes past g compute unused values
#pragma unroll
    for (int a = 0; a < AOV_LMAX_NA; ++a) {
      const double* t = Tz + ((long)iz * na + min(a, na - 1)) * nl;
      tz[zc][a][0] = t[0];
      tz[zc][a][1] = nl > 1 ? t[1] : 0.0;         // power 1: p functions only
    }
  }
  T* out = phi + ((long)b * nao_pad + (long)A * nloc) * ld;
  const int l1 = min(g * g, (int)(blockIdx.x + 1) * AOV_LINES);
  for (int line = blockIdx.x * AOV_LINES + w; line < l1; line += AOV_LNW) {
    const int ix = line / g, iy = line - ix * g;
    const double* px_t = Tx + (long)ix * na * nl;
    const double* py_t = Ty + (long)iy * na * nl;
    const long r0 = (long)line * g;
    for (int o = 0; o < nloc; ++o) {
      const int c = comp[o];
      const int px = (c == 0), py = (c == 1), pz = (c == 2);
      double acc[ZC];
#pragma unroll
      for (int zc = 0; zc < ZC; ++zc) acc[zc] = 0.0;
#pragma unroll
      for (int a = 0; a < AOV_LMAX_NA; ++a) {
        if (a >= na) break;                          // wave-uniform
        const double wa = coef[a * nloc + o];
        if (wa == 0.0) continue;
        const double cxy = wa * px_t[a * nl + px] * py_t[a * nl + py];
#pragma unroll
        for (int zc = 0; zc < ZC; ++zc) acc[zc] += cxy * (pz ? tz[zc][a][1] : tz[zc][a][0]);
      }
#pragma unroll
      for (int zc = 0; zc < ZC; ++zc) {
        const int iz = lane + 64 * zc;
        if (iz < g) {
          out[(long)o * ld + r0 + iz] = (T)acc[zc];
          mx = fmax(mx, fabs(acc[zc]));
        }
      }
    }
  }
  if (amax) {
    // the workgroup's maximum by one 64-bit atomic max on its bit pattern
    // (non-negative doubles order as their bits)
    const double m = block_max(mx, red);
    if (threadIdx.x == 0) atomicMax(amax + b, (unsigned long long)__double_as_longlong(m));
  }
}

template <typename T>
static int launch_ao_values(const double* P, int B, int n, int g, int na, int nl,
                            const double* coef, const int* comp, int nloc, int nao_pad, int ld,
                            T* phi, void* stream, double* amax_d = nullptr) {
  unsigned long long* amax = reinterpret_cast<unsigned long long*>(amax_d);
  if (B < 0 || n <= 0 || g <= 0 || na <= 0 || nl <= 0 || nloc <= 0) return HQC_E_ARG;
  if (na > HQC_MAX_NA || nloc > HQC_MAX_NLOC) return HQC_E_SIZE;
  if (nao_pad < n * nloc || ld < g * g * g) return HQC_E_ARG;
  if (B == 0) return HQC_OK;
  const int N = g * g * g;
  if (g <= 128 && na <= AOV_LMAX_NA && !hqc_tuned(HQC_TUNE_AO_POINTWISE, 0)) {
    const dim3 grid(hqc_div_up(g * g, AOV_LINES), n, B);
    if (g <= 64)
      hipLaunchKernelGGL((k_ao_values_lines<T, 1>), grid, dim3(64 * AOV_LNW), 0,
                         (hipStream_t)stream, P, n, g, na, nl, coef, comp, nloc, nao_pad, ld, phi,
                         amax);
    else
      hipLaunchKernelGGL((k_ao_values_lines<T, 2>), grid, dim3(64 * AOV_LNW), 0,
                         (hipStream_t)stream, P, n, g, na, nl, coef, comp, nloc, nao_pad, ld, phi,
                         amax);
    HQC_CHECK_LAUNCH();
    return HQC_OK;
  }
  hipLaunchKernelGGL(k_ao_values<T>, dim3(hqc_div_up(N, 256 * AOV_PPT), n, B), dim3(256), 0,
                     (hipStream_t)stream, P, n, g, na, nl, coef, comp, nloc, nao_pad, ld, phi,
                     amax);
  HQC_CHECK_LAUNCH();
  return HQC_OK;
}

extern "C" int hqc_ao_values_f64(const double* P, int B, int n, int g, int na, int nl,
                                 const double* coef, const int* comp, int nloc, int nao_pad,
                                 int ld, double* phi, void* stream) {
  return launch_ao_values<double>(P, B, n, g, na, nl, coef, comp, nloc, nao_pad, ld, phi,
                                  stream);
}

extern "C" int hqc_ao_values_f64_amax(const double* P, int B, int n, int g, int na, int nl,
                                      const double* coef, const int* comp, int nloc, int nao_pad,
                                      int ld, double* phi, double* amax, void* stream) {
  if (!amax) return HQC_E_ARG;
  return launch_ao_values<double>(P, B, n, g, na, nl, coef, comp, nloc, nao_pad, ld, phi,
                                  stream, amax);
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

// ---------------------------------------------------------------------------
// Grid adjoint, line form.  Each primitive a of an AO is a product of 1-D
// factors, so with Z = 2 v X (+ Zx) of the AO's row
//   sum_r Z(r) d/dx phi(r) = sum_a w_a sum_iz t_z(iz) sum_{ix,iy} t'_x(ix) t_y(iy) Z(ix,iy,iz)
// (t' the derivative factor; y and z alike).  Lane = iz, so a grid line (ix, iy)
// is one coalesced wave load; the line's x and y factors are the same for every
// lane (LDS broadcasts), and each lane keeps three running sums per primitive.
// The z factors and the reduction over lanes and waves come once at the end.
// Per point and primitive: three FMAs against the pointwise form's table reads
// and products in all three directions.
// Work split: the HQC_FORCE_BLOCKS slots of an atom are C = HQC_FORCE_BLOCKS /
// nloc contiguous line ranges of each AO row (slot = o C + chunk; slots past
// nloc C get zeros), and a workgroup's waves take U consecutive lines each in
// turn: workgroups that run together read neighbouring parts of the same rows.
// (One workgroup per whole row ran at a third of the copy rate, its thousands of
// concurrent streams each at another place in HBM.)  ZC = ceil(g / 64) lane
// chunks; GFL_PCH primitives per pass (one pass for these bases).
// ---------------------------------------------------------------------------
#define GFL_NW 4
#define GFL_PCH 4
#define GFL_MAXG 128

// MODE bit 0: v (and X) given; bit 1: Zx given.  Every load of a batch of U lines
// is issued before the first is used (no branch between them: lines past the
// range and lanes past the line are clamped and masked to zero).
template <int ZC, int MODE, int U>
__global__ __launch_bounds__(64 * GFL_NW) void k_grid_forces_lines(
    const double* __restrict__ P, int n, int g, int na, int nl,
    const double* __restrict__ alpha, const double* __restrict__ coef,
    const int* __restrict__ comp, int nloc, int nao_pad, int ld,
    const double* __restrict__ X, const double* __restrict__ v,
    const double* __restrict__ Zx, double dV, double* __restrict__ part) {
  __shared__ double red[4];
  __shared__ double2 sx[GFL_MAXG][GFL_PCH];       // (value, derivative) factors per ix
  __shared__ double2 sy[GFL_MAXG][GFL_PCH];       // the same per iy
  const int slot = blockIdx.x, A = blockIdx.y, b = blockIdx.z;
  double* out = part + (((long)b * n + A) * gridDim.x + slot) * 3;
  const int C = gridDim.x / nloc;                 // line ranges per AO
  const int o = slot / C, chunk = slot - o * C;
  if (o >= nloc) {                               // unused slots: zero partials
    if (threadIdx.x < 3) out[threadIdx.x] = 0.0;
    return;
  }
  const int nline = g * g;
  const int per = (nline + C - 1) / C;
  const int lb = chunk * per, le = min(nline, lb + per);
  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int c = comp[o];
  const int px = (c == 0), py = (c == 1), pz = (c == 2);
  const long tstride = (long)g * na * nl;         // one direction's tables
  const double* Tx = P + (long)(b * n + A) * 3 * tstride;
  const double* Ty = Tx + tstride;
  const double* Tz = Ty + tstride;
  const long row = (long)b * nao_pad + (long)A * nloc + o;
  const double* Xr = (MODE & 1) ? X + row * ld : nullptr;    // read only as MODE says
  const double* Zr = (MODE & 2) ? Zx + row * ld : nullptr;
  const double* vb = (MODE & 1) ? v + (long)b * ld : nullptr;
  double f[3] = {0.0, 0.0, 0.0};
  for (int a0 = 0; a0 < na && lb < le;) {
    // the next GFL_PCH primitives with a nonzero coefficient in this AO
    int act[GFL_PCH];
    double wt[GFL_PCH], m2a[GFL_PCH];
    int np = 0, a = a0;
#pragma unroll
    for (int j = 0; j < GFL_PCH; ++j) {
      while (a < na && coef[a * nloc + o] == 0.0) ++a;
      act[j] = a < na ? a : 0;
      wt[j] = a < na ? coef[a * nloc + o] : 0.0;
      m2a[j] = -2.0 * alpha[act[j]];
      if (a < na) {
        ++np;
        ++a;
      }
    }
    a0 = a;
    if (np == 0) break;
    __syncthreads();                              // the previous chunk's factors are read
    for (int i = threadIdx.x; i < 2 * g * GFL_PCH; i += blockDim.x) {
      const int d = i / (g * GFL_PCH), q = i - d * g * GFL_PCH;
      const int ii = q / GFL_PCH, j = q - ii * GFL_PCH;
      const int p = d ? py : px;
      int aj = act[0];
      double mj = m2a[0];
#pragma unroll
      for (int jj = 1; jj < GFL_PCH; ++jj)
        if (j == jj) {
          aj = act[jj];
          mj = m2a[jj];
        }
      const double* t = (d ? Ty : Tx) + ((long)ii * na + aj) * nl;
      const double2 fv = make_double2(t[p], mj * t[p + 1] + (p ? t[0] : 0.0));
      if (d) sy[ii][j] = fv; else sx[ii][j] = fv;
    }
    __syncthreads();
    double acc[ZC][3][GFL_PCH];
#pragma unroll
    for (int zc = 0; zc < ZC; ++zc)
#pragma unroll
      for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int j = 0; j < GFL_PCH; ++j) acc[zc][k][j] = 0.0;
    for (int l0 = lb + w * U; l0 < le; l0 += GFL_NW * U) {
      double z[U][ZC];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const long r0 = (long)min(l0 + u, le - 1) * g;
#pragma unroll
        for (int zc = 0; zc < ZC; ++zc) {
          const long r = r0 + min(lane + 64 * zc, g - 1);
          double zz = 0.0;
          if (MODE & 1) zz = 2.0 * vb[r] * Xr[r];
          if (MODE & 2) zz += Zr[r];
          z[u][zc] = zz;
        }
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int l = min(l0 + u, le - 1);
        const int ix = l / g, iy = l - ix * g;
        const bool line_ok = l0 + u < le;         // wave-uniform
#pragma unroll
        for (int zc = 0; zc < ZC; ++zc)
          z[u][zc] = (line_ok && lane + 64 * zc < g) ? z[u][zc] : 0.0;
#pragma unroll
        for (int j = 0; j < GFL_PCH; ++j) {
          if (j >= np) continue;                   // wave-uniform
          const double2 xf = sx[ix][j], yf = sy[iy][j];
          const double c1 = xf.y * yf.x, c2 = xf.x * yf.y, c3 = xf.x * yf.x;
#pragma unroll
          for (int zc = 0; zc < ZC; ++zc) {
            acc[zc][0][j] = fma(c1, z[u][zc], acc[zc][0][j]);
            acc[zc][1][j] = fma(c2, z[u][zc], acc[zc][1][j]);
            acc[zc][2][j] = fma(c3, z[u][zc], acc[zc][2][j]);
          }
        }
      }
    }
#pragma unroll
    for (int zc = 0; zc < ZC; ++zc) {
      const int iz = lane + 64 * zc;
      if (iz >= g) continue;
#pragma unroll
      for (int j = 0; j < GFL_PCH; ++j) {
        if (j >= np) continue;
        const double* t = Tz + ((long)iz * na + act[j]) * nl;
        const double vz = t[pz];
        const double dz = m2a[j] * t[pz + 1] + (pz ? t[0] : 0.0);
        f[0] = fma(wt[j] * vz, acc[zc][0][j], f[0]);
        f[1] = fma(wt[j] * vz, acc[zc][1][j], f[1]);
        f[2] = fma(wt[j] * dz, acc[zc][2][j], f[2]);
      }
    }
  }
  for (int k = 0; k < 3; ++k) {
    const double t = block_sum(f[k], red);
    if (threadIdx.x == 0) out[k] = t * dV;
  }
}

template <int ZC, int U>
static void launch_grid_forces_lines(dim3 grid, hipStream_t s, const double* P, int n, int g,
                                     int na, int nl, const double* alpha, const double* coef,
                                     const int* comp, int nloc, int nao_pad, int ld,
                                     const double* X, const double* v, const double* Zx,
                                     double dV, double* part) {
  const int mode = (v ? 1 : 0) | (Zx ? 2 : 0);
#define GFL_LAUNCH(M)                                                                           \
  hipLaunchKernelGGL((k_grid_forces_lines<ZC, M, U>), grid, dim3(64 * GFL_NW), 0, s, P, n, g, \
                     na, nl, alpha, coef, comp, nloc, nao_pad, ld, X, v, Zx, dV, part)
  if (mode == 1) GFL_LAUNCH(1);
  else if (mode == 2) GFL_LAUNCH(2);
  else if (mode == 3) GFL_LAUNCH(3);
  else GFL_LAUNCH(0);
#undef GFL_LAUNCH
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
  if (nloc > HQC_FORCE_BLOCKS) return HQC_E_SIZE;
  if (g <= GFL_MAXG && !hqc_tuned(HQC_TUNE_FORCE_POINTWISE, 0)) {
    // line form (one workgroup per AO; slots o >= nloc get zeros)
    const dim3 grid(HQC_FORCE_BLOCKS, n, B);
    if (g <= 64)
      launch_grid_forces_lines<1, 8>(grid, (hipStream_t)stream, P, n, g, na, nl, alpha, coef,
                                     comp, nloc, nao_pad, ld, X, v, Zx, dV, part);
    else
      launch_grid_forces_lines<2, 4>(grid, (hipStream_t)stream, P, n, g, na, nl, alpha, coef,
                                     comp, nloc, nao_pad, ld, X, v, Zx, dV, part);
    HQC_CHECK_LAUNCH();
    return HQC_OK;
  }
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
