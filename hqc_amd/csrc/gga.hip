// GGA (PBE) exchange-correlation on the grid with spectral derivatives.
//
// Not in the reference, which accepts only LDA exchange
// (hqc/tools/excor.py:14-17).  The discrete functional is
//   E_xc = dV sum_r e(rho_r, sigma_r),  sigma = |D rho|^2,
// with D the FFT derivative on the odd grid (real, no Nyquist term), and its
// exact derivative with respect to the grid values is
//   v_xc = de/drho - D . (2 de/dsigma D rho)       (D^T = -D),
// so the potential matrix and the force adjoint are those of a local
// potential.  e is PBE (Perdew, Burke, Ernzerhof, PRL 77, 3865 (1996)),
// unpolarised, with PW92 LDA correlation (PRB 45, 13244, eq. 10); both
// partial derivatives come from forward-mode dual numbers.
#include "common.h"

#define GGA_BLOCKS 256
#define GGA_RHO_CUT 1e-12

namespace {

// value and partial derivatives with respect to (rho, sigma)
struct dual2 {
  double v, r, s;
};

__device__ __forceinline__ dual2 mk(double v) { return {v, 0.0, 0.0}; }
__device__ __forceinline__ dual2 operator+(dual2 a, dual2 b) { return {a.v + b.v, a.r + b.r, a.s + b.s}; }
__device__ __forceinline__ dual2 operator-(dual2 a, dual2 b) { return {a.v - b.v, a.r - b.r, a.s - b.s}; }
__device__ __forceinline__ dual2 operator*(dual2 a, dual2 b) {
  return {a.v * b.v, a.r * b.v + a.v * b.r, a.s * b.v + a.v * b.s};
}
__device__ __forceinline__ dual2 operator/(dual2 a, dual2 b) {
  const double q = a.v / b.v;
  return {q, (a.r - q * b.r) / b.v, (a.s - q * b.s) / b.v};
}
__device__ __forceinline__ dual2 operator+(dual2 a, double c) { return {a.v + c, a.r, a.s}; }
__device__ __forceinline__ dual2 operator*(double c, dual2 a) { return {c * a.v, c * a.r, c * a.s}; }
__device__ __forceinline__ dual2 operator-(double c, dual2 a) { return {c - a.v, -a.r, -a.s}; }
__device__ __forceinline__ dual2 operator/(double c, dual2 b) {
  const double q = c / b.v;
  return {q, -q * b.r / b.v, -q * b.s / b.v};
}
// f(a) with f' given
__device__ __forceinline__ dual2 chain(dual2 a, double f, double df) {
  return {f, df * a.r, df * a.s};
}
__device__ __forceinline__ dual2 dlog(dual2 a) { return chain(a, log(a.v), 1.0 / a.v); }
__device__ __forceinline__ dual2 dexp(dual2 a) {
  const double e = exp(a.v);
  return chain(a, e, e);
}
__device__ __forceinline__ dual2 dsqrt(dual2 a) {
  const double q = sqrt(a.v);
  return chain(a, q, 0.5 / q);
}
__device__ __forceinline__ dual2 dcbrt(dual2 a) {
  const double q = cbrt(a.v);
  return chain(a, q, q / (3.0 * a.v));
}

// e(rho, sigma) per volume
__device__ dual2 pbe_e(double rho_v, double sigma_v) {
  const double kappa = 0.804, beta = 0.06672455060314922;
  const double mu = beta * M_PI * M_PI / 3.0;
  const double gamma = (1.0 - M_LN2) / (M_PI * M_PI);
  const dual2 rho = {rho_v, 1.0, 0.0}, sigma = {sigma_v, 0.0, 1.0};
  const dual2 r13 = dcbrt(rho);
  const dual2 kF = cbrt(3.0 * M_PI * M_PI) * r13;
  const dual2 rho2 = rho * rho;
  // exchange
  const dual2 ex_lda = (-0.75 * cbrt(3.0 / M_PI)) * r13;
  const dual2 s2 = sigma / (4.0 * (kF * kF * rho2));
  const dual2 Fx = (1.0 + kappa) - kappa / (mk(1.0) + (mu / kappa) * s2);
  // PW92 correlation (unpolarised)
  const double A = 0.0310907, a1 = 0.21370, b1 = 7.5957, b2 = 3.5876, b3 = 1.6382,
               b4 = 0.49294;
  const dual2 rs = cbrt(3.0 / (4.0 * M_PI)) / r13;
  const dual2 srs = dsqrt(rs);
  const dual2 q = (2.0 * A) * (b1 * srs + b2 * rs + b3 * (rs * srs) + b4 * (rs * rs));
  const dual2 ec = (-2.0 * A) * ((mk(1.0) + a1 * rs) * dlog(mk(1.0) + mk(1.0) / q));
  // gradient correction H(t), t^2 = sigma / (2 k_s rho)^2, k_s = sqrt(4 k_F / pi)
  const dual2 ks2 = (4.0 / M_PI) * kF;
  const dual2 t2 = sigma / (4.0 * (ks2 * rho2));
  const double bg = beta / gamma;
  const dual2 Ac = mk(bg) / (dexp((-1.0 / gamma) * ec) + (-1.0));
  const dual2 At2 = Ac * t2;
  const dual2 H = gamma * dlog(mk(1.0) + (bg * t2) * (mk(1.0) + At2) / (mk(1.0) + At2 + At2 * At2));
  return rho * (ex_lda * Fx + ec + H);
}

__device__ __forceinline__ double gvec(int k, int g, double L) {
  const int kf = (k <= g / 2) ? k : k - g;
  return (double)kf * 2.0 * M_PI / L;
}

// out[b, d, k] = i G_d(k) rhoG[b, k] * scale
__global__ __launch_bounds__(256) void k_grad_G(const double2* __restrict__ rg, int g, double L0,
                                                const double* __restrict__ Lb, double scale,
                                                double2* __restrict__ out) {
  const int N = g * g * g;
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y;
  if (k >= N) return;
  const double L = Lb ? Lb[b] : L0;
  const int kz = k % g, ky = (k / g) % g, kx = k / (g * g);
  const double G[3] = {gvec(kx, g, L), gvec(ky, g, L), gvec(kz, g, L)};
  const double2 z = rg[(long)b * N + k];
#pragma unroll
  for (int d = 0; d < 3; ++d)
    out[((long)b * 3 + d) * N + k] = make_double2(-G[d] * z.y * scale, G[d] * z.x * scale);
}

// out[b, k] = scale * sum_d i G_d(k) wG[b, d, k]
__global__ __launch_bounds__(256) void k_div_G(const double2* __restrict__ wg, int g, double L0,
                                               const double* __restrict__ Lb, double scale,
                                               double2* __restrict__ out) {
  const int N = g * g * g;
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y;
  if (k >= N) return;
  const double L = Lb ? Lb[b] : L0;
  const int kz = k % g, ky = (k / g) % g, kx = k / (g * g);
  const double G[3] = {gvec(kx, g, L), gvec(ky, g, L), gvec(kz, g, L)};
  double re = 0.0, im = 0.0;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const double2 z = wg[((long)b * 3 + d) * N + k];
    re -= G[d] * z.y;
    im += G[d] * z.x;
  }
  out[(long)b * N + k] = make_double2(re * scale, im * scale);
}

// vout = v_H + de/drho, w_d = 2 de/dsigma (D_d rho) (complex, zero imaginary part),
// per-block partial sums of e and rho v_H
__global__ __launch_bounds__(256) void k_xc_gga(const double* __restrict__ rho,
                                                const double2* __restrict__ grad,
                                                const double2* __restrict__ vh, int N, int ld,
                                                double* __restrict__ vout,
                                                double2* __restrict__ w,
                                                double* __restrict__ part) {
  __shared__ double red[4];
  const int b = blockIdx.y;
  double se = 0.0, sh = 0.0;
  for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < N; r += gridDim.x * blockDim.x) {
    const double p = rho[(long)b * ld + r];
    const double h = vh[(long)b * N + r].x;
    const double g0 = grad[((long)b * 3 + 0) * N + r].x;
    const double g1 = grad[((long)b * 3 + 1) * N + r].x;
    const double g2 = grad[((long)b * 3 + 2) * N + r].x;
    double e = 0.0, vr = 0.0, vs = 0.0;
    if (p > GGA_RHO_CUT) {
      const dual2 E = pbe_e(p, g0 * g0 + g1 * g1 + g2 * g2);
      e = E.v;
      vr = E.r;
      vs = E.s;
    }
    vout[(long)b * ld + r] = h + vr;
    w[((long)b * 3 + 0) * N + r] = make_double2(2.0 * vs * g0, 0.0);
    w[((long)b * 3 + 1) * N + r] = make_double2(2.0 * vs * g1, 0.0);
    w[((long)b * 3 + 2) * N + r] = make_double2(2.0 * vs * g2, 0.0);
    se += e;
    sh += p * h;
  }
  se = block_sum(se, red);
  sh = block_sum(sh, red);
  if (threadIdx.x == 0) {
    part[((long)b * gridDim.x + blockIdx.x) * 2] = se;
    part[((long)b * gridDim.x + blockIdx.x) * 2 + 1] = sh;
  }
}

// point values for tests: e, de/drho, de/dsigma
__global__ void k_pbe_points(const double* __restrict__ rho, const double* __restrict__ sigma,
                             int n, double* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double e = 0.0, vr = 0.0, vs = 0.0;
  if (rho[i] > GGA_RHO_CUT) {
    const dual2 E = pbe_e(rho[i], sigma[i]);
    e = E.v;
    vr = E.r;
    vs = E.s;
  }
  out[3 * i] = e;
  out[3 * i + 1] = vr;
  out[3 * i + 2] = vs;
}

}  // namespace

extern "C" int hqc_grad_G(const double* rhoG, int B, int g, double L, const double* Lb,
                          double scale, double* out, void* stream) {
  if (B < 0 || g <= 0 || !(L > 0.0)) return HQC_E_ARG;
  if (B == 0) return HQC_OK;
  const int N = g * g * g;
  hipLaunchKernelGGL(k_grad_G, dim3(hqc_div_up(N, 256), B), dim3(256), 0, (hipStream_t)stream,
                     (const double2*)rhoG, g, L, Lb, scale, (double2*)out);
  HQC_CHECK_LAUNCH();
  return HQC_OK;
}

extern "C" int hqc_div_G(const double* wG, int B, int g, double L, const double* Lb,
                         double scale, double* out, void* stream) {
  if (B < 0 || g <= 0 || !(L > 0.0)) return HQC_E_ARG;
  if (B == 0) return HQC_OK;
  const int N = g * g * g;
  hipLaunchKernelGGL(k_div_G, dim3(hqc_div_up(N, 256), B), dim3(256), 0, (hipStream_t)stream,
                     (const double2*)wG, g, L, Lb, scale, (double2*)out);
  HQC_CHECK_LAUNCH();
  return HQC_OK;
}

extern "C" int hqc_xc_gga_eval(const double* rho, const double* grad, const double* vh, int B,
                               int g, int ld, double* vout, double* w, double* part,
                               void* stream) {
  if (B < 0 || g <= 0 || ld < g * g * g) return HQC_E_ARG;
  if (B == 0) return HQC_OK;
  hipLaunchKernelGGL(k_xc_gga, dim3(GGA_BLOCKS, B), dim3(256), 0, (hipStream_t)stream, rho,
                     (const double2*)grad, (const double2*)vh, g * g * g, ld, vout, (double2*)w,
                     part);
  HQC_CHECK_LAUNCH();
  return HQC_OK;
}

extern "C" int hqc_gga_blocks(int g) {
  (void)g;
  return GGA_BLOCKS;
}

extern "C" int hqc_pbe_points(const double* rho, const double* sigma, int n, double* out,
                              void* stream) {
  if (n < 0) return HQC_E_ARG;
  if (n == 0) return HQC_OK;
  hipLaunchKernelGGL(k_pbe_points, dim3(hqc_div_up(n, 256)), dim3(256), 0, (hipStream_t)stream,
                     rho, sigma, n, out);
  HQC_CHECK_LAUNCH();
  return HQC_OK;
}
