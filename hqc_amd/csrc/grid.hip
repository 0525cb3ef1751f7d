// Pointwise grid kernels: local potential in G space, Poisson multiply,
// exchange-correlation, structure-factor force, plus the small per-geometry
// kernels (occupations / chemical potential, Ewald pair sums).
#include "common.h"

// G vector component for FFT index k (FFT order), formed as in
// hqc/pbc/solver.py:195: (k*2)*pi/L.
__device__ __forceinline__ double gcomp(int k, int g, double L) {
  const int kf = (k <= g / 2) ? k : k - g;
  return (double)kf * 2.0 * M_PI / L;
}

// ---------------------------------------------------------------------------
// Structure factors exp(-i G.R_A) on the odd grid factor into 1-D phases,
// exp(-i Gx x) exp(-i Gy y) exp(-i Gz z): a workgroup builds the x, y and z
// phase tables of a chunk of atoms in LDS (3 g sincos per atom instead of one
// per grid point and atom) and forms each point's phase from two complex
// products.  Atoms in chunks of at most SF_ATOMS, fewer where the tables (chunk x
// 3 g x 16 bytes) would pass 64 KB.
// ---------------------------------------------------------------------------
#define SF_ATOMS 16
#define SF_PTS 8            // grid points per thread in k_vloc_G

// tab[a][d][k] = exp(i sgn G_d(k) x_{a,d}) for the atoms [a0, a0 + na) of x (n x 3)
__device__ __forceinline__ void sf_tables(const double* __restrict__ x, int a0, int na, int g,
                                          double L, double sgn, double2* tab) {
  for (int t = threadIdx.x; t < na * 3 * g; t += blockDim.x) {
    const int k = t % g, d = (t / g) % 3, a = t / (3 * g);
    double s, c;
    sincos(sgn * gcomp(k, g, L) * x[(a0 + a) * 3 + d], &s, &c);
    tab[t] = make_double2(c, s);
  }
}

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// v_loc(G) = -VG sum_A exp(-i G.R_A)     (hqc/pbc/solver.py:406-407)
// grid (ceil(N / (256 SF_PTS)), B): SF_PTS points per thread, strided by 256
__global__ __launch_bounds__(256) void k_vloc_G(const double* __restrict__ xp, int n, int g,
                                                double L, int at, const double* __restrict__ VG,
                                                double2* __restrict__ out) {
  extern __shared__ double2 tab[];                // [at][3][g]
  const int N = g * g * g;
  const int b = blockIdx.y;
  const int r0 = blockIdx.x * 256 * SF_PTS + threadIdx.x;
  const double* x = xp + (long)b * n * 3;
  double2 acc[SF_PTS];
  int kx[SF_PTS], ky[SF_PTS], kz[SF_PTS];
#pragma unroll
  for (int j = 0; j < SF_PTS; ++j) {
    acc[j] = make_double2(0.0, 0.0);
    const int r = min(r0 + 256 * j, N - 1);
    kz[j] = r % g;
    ky[j] = (r / g) % g;
    kx[j] = r / (g * g);
  }
  for (int a0 = 0; a0 < n; a0 += at) {
    const int na = min(at, n - a0);
    __syncthreads();                              // the previous chunk's tables are read
    sf_tables(x, a0, na, g, L, -1.0, tab);
    __syncthreads();
    for (int a = 0; a < na; ++a) {
      const double2* ta = tab + a * 3 * g;
#pragma unroll
      for (int j = 0; j < SF_PTS; ++j) {
        const double2 p = cmul(cmul(ta[kx[j]], ta[g + ky[j]]), ta[2 * g + kz[j]]);
        acc[j].x += p.x;
        acc[j].y += p.y;
      }
    }
  }
#pragma unroll
  for (int j = 0; j < SF_PTS; ++j) {
    const int r = r0 + 256 * j;
    if (r < N) {
      const double vg = VG[r];
      out[(long)b * N + r] = make_double2(-acc[j].x * vg, -acc[j].y * vg);
    }
  }
}

// atoms per table chunk: SF_ATOMS, or fewer so the chunk's tables fit 64 KB
static int sf_chunk(int g) {
  const int at = (int)(65536 / (3 * (long)g * sizeof(double2)));
  return at < SF_ATOMS ? at : SF_ATOMS;
}

extern "C" int hqc_vloc_G(const double* xp, int B, int n, int g, double L, const double* VG,
                          double* out, void* stream) {
  if (B < 0 || n <= 0 || g <= 0 || sf_chunk(g) < 1) return HQC_E_ARG;
  if (B == 0) return HQC_OK;
  const int N = g * g * g;
  const int at = sf_chunk(g);
  const size_t lds = (size_t)at * 3 * g * sizeof(double2);
  hipLaunchKernelGGL(k_vloc_G, dim3(hqc_div_up(N, 256 * SF_PTS), B), dim3(256), lds,
                     (hipStream_t)stream, xp, n, g, L, at, VG, (double2*)out);
  HQC_CHECK_LAUNCH();
  return HQC_OK;
}

// ---------------------------------------------------------------------------
// Layout conversions and the Poisson multiply.
// ---------------------------------------------------------------------------
__global__ void k_r2c(const double* __restrict__ in, int N, int ld, double s,
                      double2* __restrict__ out) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y;
  if (r >= N) return;
  out[(long)b * N + r] = make_double2(s * in[(long)b * ld + r], 0.0);
}

// two real grids per complex grid: out[p] = s (in[2p] + i in[2p+1]) (the last
// imaginary part 0 for odd B); with a real, even kernel such as VG the transform,
// multiply and inverse transform keep the two grids apart in Re and Im
__global__ void k_r2c_pair(const double* __restrict__ in, int B, int N, int ld, double s,
                           double2* __restrict__ out) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  const int p = blockIdx.y;
  if (r >= N) return;
  const double re = s * in[(long)(2 * p) * ld + r];
  const double im = 2 * p + 1 < B ? s * in[(long)(2 * p + 1) * ld + r] : 0.0;
  out[(long)p * N + r] = make_double2(re, im);
}

__global__ void k_c2r(const double2* __restrict__ in, int N, int ld, double* __restrict__ out) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y;
  if (r >= N) return;
  out[(long)b * ld + r] = in[(long)b * N + r].x;
}

__global__ void k_poisson_mul(double2* __restrict__ c, const double* __restrict__ VG, int N,
                              long vg_stride, double2* __restrict__ saved) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y;
  if (r >= N) return;
  const long i = (long)b * N + r;
  double2 v = c[i];
  if (saved) saved[i] = v;
  const double w = VG[b * vg_stride + r];
  c[i] = make_double2(v.x * w, v.y * w);
}

extern "C" int hqc_real_to_complex(const double* in, int B, int g, int ld, double s, double* out,
                                   void* stream) {
  if (B < 0 || g <= 0 || ld < g * g * g) return HQC_E_ARG;
  if (B == 0) return HQC_OK;
  const int N = g * g * g;
  hipLaunchKernelGGL(k_r2c, dim3(hqc_div_up(N, 256), B), dim3(256), 0, (hipStream_t)stream, in, N,
                     ld, s, (double2*)out);
  HQC_CHECK_LAUNCH();
  return HQC_OK;
}

extern "C" int hqc_real_pair_to_complex(const double* in, int B, int g, int ld, double s,
                                        double* out, void* stream) {
  if (B < 0 || g <= 0 || ld < g * g * g) return HQC_E_ARG;
  if (B == 0) return HQC_OK;
  const int N = g * g * g;
  hipLaunchKernelGGL(k_r2c_pair, dim3(hqc_div_up(N, 256), (B + 1) / 2), dim3(256), 0,
                     (hipStream_t)stream, in, B, N, ld, s, (double2*)out);
  HQC_CHECK_LAUNCH();
  return HQC_OK;
}

extern "C" int hqc_complex_real(const double* in, int B, int g, int ld, double* out,
                                void* stream) {
  if (B < 0 || g <= 0 || ld < g * g * g) return HQC_E_ARG;
  if (B == 0) return HQC_OK;
  const int N = g * g * g;
  hipLaunchKernelGGL(k_c2r, dim3(hqc_div_up(N, 256), B), dim3(256), 0, (hipStream_t)stream,
                     (const double2*)in, N, ld, out);
  HQC_CHECK_LAUNCH();
  return HQC_OK;
}

extern "C" int hqc_poisson_mul(double* c, const double* VG, int B, int g, int vg_batched,
                               double* saved, void* stream) {
  if (B < 0 || g <= 0) return HQC_E_ARG;
  if (B == 0) return HQC_OK;
  const int N = g * g * g;
  hipLaunchKernelGGL(k_poisson_mul, dim3(hqc_div_up(N, 256), B), dim3(256), 0,
                     (hipStream_t)stream, (double2*)c, VG, N, vg_batched ? (long)N : 0L,
                     (double2*)saved);
  HQC_CHECK_LAUNCH();
  return HQC_OK;
}

// ---------------------------------------------------------------------------
// Exchange-correlation (hqc/tools/excor.py) with closed-form derivatives:
//   v_xc = eps + rho d eps/d rho   (the reference takes this with jax.grad,
//   hqc/pbc/solver.py:673).
// ---------------------------------------------------------------------------
// workgroups per geometry: enough points per thread to hide the FP64
// transcendentals' latency at full batches (8192 workgroups in all), at most 256
// for a few geometries (B = 512: 0.83 against 1.17 ms with 256 per geometry)
#define XC_TARGET_WGS 8192
#define XC_MIN_BLOCKS 16
#define XC_MAX_BLOCKS 256

__device__ __forceinline__ void xc_point(double rho, int xc_id, double& eps, double& v) {
  const double r13 = cbrt(rho);
  const double ex = -0.75 * cbrt(3.0 / M_PI) * r13;   // -3/4 (3/pi)^(1/3) rho^(1/3)
  eps = ex;
  v = 4.0 / 3.0 * ex;
  if (xc_id == 2) return;
  const double ws = cbrt(3.0 / 4.0 / M_PI);
  const double rs = ws / r13;
  if (xc_id == 1) {
    // VWN parameters of hqc/tools/excor.py:31-43
    const double A = 0.0621814, y0 = -0.10498, bb = 3.72744, cc = 12.93532;
    const double Yy0 = y0 * y0 + bb * y0 + cc;
    const double Q = sqrt(4.0 * cc - bb * bb);
    const double c0 = A / 2.0;
    const double c1 = c0 * 2.0 * bb / Q;
    const double c2 = c0 * bb * y0 / Yy0;
    const double c3 = c2 * 2.0 * (bb + 2.0 * y0) / Q;
    const double c4 = c1 - c3;
    const double y = sqrt(rs);
    const double Y = y * y + bb * y + cc;
    const double ec = c0 * log(y * y / Y) + c4 * atan(Q / (2.0 * y + bb)) -
                      c2 * log((y - y0) * (y - y0) / Y);
    const double dY = 2.0 * y + bb;
    const double dec = c0 * (2.0 / y - dY / Y) - c4 * Q / (2.0 * Y) - c2 * (2.0 / (y - y0) - dY / Y);
    eps += ec;
    v += ec - y / 6.0 * dec;
    return;
  }
  double Ac, Cc;
  if (xc_id == 3) {
    Ac = 21.0; Cc = 0.045;        // HL71
  } else if (xc_id == 4) {
    Ac = 30.0; Cc = 0.0504;       // BH72
  } else {
    Ac = 11.4; Cc = 0.0666;       // GL77
  }
  const double x = rs / Ac;
  const double lg = log(1.0 + 1.0 / x);
  const double ec = -0.5 * Cc * ((1.0 + x * x * x) * lg + x / 2.0 - x * x - 1.0 / 3.0);
  const double dec = -0.5 * Cc * (3.0 * x * x * lg - 1.0 / x + 1.5 - 3.0 * x);
  eps += ec;
  v += ec - x / 3.0 * dec;
}

__global__ __launch_bounds__(256) void k_xc_eval(const double* __restrict__ rho,
                                                 const double2* __restrict__ vh, int N, int ld,
                                                 int xc_id, int paired,
                                                 double* __restrict__ vout,
                                                 double* __restrict__ part) {
  __shared__ double red[4];
  const int b = blockIdx.y;
  // paired: geometry b's v_H is Re (b even) or Im (b odd) of complex grid b/2
  const double* vhb = reinterpret_cast<const double*>(vh) +
                      (paired ? (long)(b >> 1) * N * 2 + (b & 1) : (long)b * N * 2);
  double se = 0.0, sh = 0.0;
  for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < N; r += gridDim.x * blockDim.x) {
    const double p = rho[(long)b * ld + r];
    const double h = vhb[2 * (long)r];
    double eps = 0.0, v = 0.0;
    if (xc_id != 0) xc_point(p, xc_id, eps, v);
    vout[(long)b * ld + r] = h + v;
    se += p * eps;
    sh += p * h;
  }
  se = block_sum(se, red);
  sh = block_sum(sh, red);
  if (threadIdx.x == 0) {
    part[((long)b * gridDim.x + blockIdx.x) * 2] = se;
    part[((long)b * gridDim.x + blockIdx.x) * 2 + 1] = sh;
  }
}

static int xc_blocks(int B) {
  int d = XC_TARGET_WGS / (B > 0 ? B : 1);
  d = d < XC_MIN_BLOCKS ? XC_MIN_BLOCKS : (d > XC_MAX_BLOCKS ? XC_MAX_BLOCKS : d);
  const int v = hqc_tuned(HQC_TUNE_XC_BLOCKS, d);
  return v < 65535 ? v : 65535;
}

extern "C" int hqc_xc_blocks(int B, int g) {
  (void)g;
  return xc_blocks(B);
}

extern "C" int hqc_xc_eval(const double* rho, const double* vh, int vh_paired, int B, int g,
                           int ld, int xc_id, double* vout, double* part, void* stream) {
  if (B < 0 || g <= 0 || ld < g * g * g || xc_id < 0 || xc_id > 5) return HQC_E_ARG;
  if (B == 0) return HQC_OK;
  hipLaunchKernelGGL(k_xc_eval, dim3(xc_blocks(B), B), dim3(256), 0, (hipStream_t)stream, rho,
                     (const double2*)vh, g * g * g, ld, xc_id, vh_paired ? 1 : 0, vout, part);
  HQC_CHECK_LAUNCH();
  return HQC_OK;
}

// ---------------------------------------------------------------------------
// dE/dR_A of the local-potential energy through the structure factor:
//   dE[b,A,c] = Re sum_G i G_c VG e^{-iG.R_A} conj(rho~(G))
//             = sum_G G_c VG (a sin th + b cos th),  rho~ = a + i b, th = G.R_A
// ---------------------------------------------------------------------------
// grid (ceil(n / at), B): a workgroup takes a chunk of atoms over every grid
// point; with rho~ = a + i b and E_A = exp(+i G.R_A), a sin th + b cos th = Im(rho~ E_A)
__global__ __launch_bounds__(256) void k_vloc_force(const double* __restrict__ xp, int n, int g,
                                                    double L, int at,
                                                    const double* __restrict__ VG,
                                                    const double2* __restrict__ rg,
                                                    double* __restrict__ dE) {
  extern __shared__ double2 tab[];                // [at][3][g]
  __shared__ double red[4];
  const int a0 = blockIdx.x * at, b = blockIdx.y;
  const int na = min(at, n - a0);
  const int N = g * g * g;
  sf_tables(xp + (long)b * n * 3, a0, na, g, L, 1.0, tab);
  __syncthreads();
  double f[SF_ATOMS][3];
#pragma unroll
  for (int a = 0; a < SF_ATOMS; ++a) f[a][0] = f[a][1] = f[a][2] = 0.0;
  for (int r = threadIdx.x; r < N; r += blockDim.x) {
    const int kz = r % g, ky = (r / g) % g, kx = r / (g * g);
    const double2 z = rg[(long)b * N + r];
    const double vg = VG[r];
    const double wx = vg * z.x, wy = vg * z.y;
    const double Gx = gcomp(kx, g, L), Gy = gcomp(ky, g, L), Gz = gcomp(kz, g, L);
#pragma unroll
    for (int a = 0; a < SF_ATOMS; ++a) {
      if (a < na) {
        const double2* ta = tab + a * 3 * g;
        const double2 p = cmul(cmul(ta[kx], ta[g + ky]), ta[2 * g + kz]);
        const double t = wx * p.y + wy * p.x;      // Im(w E_A)
        f[a][0] += Gx * t;
        f[a][1] += Gy * t;
        f[a][2] += Gz * t;
      }
    }
  }
  // unrolled with a uniform guard: an index not known at compile time would move f
  // to scratch memory
#pragma unroll
  for (int a = 0; a < SF_ATOMS; ++a) {
    if (a < na) {
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const double v = block_sum(f[a][c], red);
        if (threadIdx.x == 0) dE[((long)b * n + a0 + a) * 3 + c] = v;
      }
    }
  }
}

extern "C" int hqc_vloc_force(const double* xp, int B, int n, int g, double L, const double* VG,
                              const double* rhoG, double* dE, void* stream) {
  if (B < 0 || n <= 0 || g <= 0 || sf_chunk(g) < 1) return HQC_E_ARG;
  if (B == 0) return HQC_OK;
  const int at = sf_chunk(g);
  const size_t lds = (size_t)at * 3 * g * sizeof(double2);
  hipLaunchKernelGGL(k_vloc_force, dim3(hqc_div_up(n, at), B), dim3(256), lds,
                     (hipStream_t)stream, xp, n, g, L, at, VG, (const double2*)rhoG, dE);
  HQC_CHECK_LAUNCH();
  return HQC_OK;
}

// ---------------------------------------------------------------------------
// Occupations (hqc/tools/smearing.py).  One 64-lane wavefront per geometry;
// every lane runs the same scalar control flow on wave-wide sums.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double occ_f(double mu, double w, double sigma, int method) {
  if (method == 0) {
    double x = (w - mu) / sigma;
    x = fmin(fmax(x, -50.0), 50.0);
    return 2.0 / (exp(x) + 1.0);
  }
  return 0.5 * (1.0 + erf((mu - w) / sigma));
}

__device__ double occ_N(const double* w, int nmo, double mu, double sigma, int method, int nelec) {
  const int lane = threadIdx.x & 63;
  double s = 0.0;
  for (int i = lane; i < nmo; i += 64) s += occ_f(mu, w[i], sigma, method);
  return wave_sum(s) - (double)nelec;
}

__device__ void occ_dN(const double* w, int nmo, double mu, double sigma, int method, double& d1,
                       double& d2) {
  const int lane = threadIdx.x & 63;
  double s1 = 0.0, s2 = 0.0;
  for (int i = lane; i < nmo; i += 64) {
    if (method == 0) {
      const double x = (w[i] - mu) / sigma;
      if (fabs(x) < 50.0) {
        const double ex = exp(x);
        const double q = ex + 1.0;
        s1 += 2.0 * ex / (q * q) / sigma;
        s2 += -2.0 * ex * (1.0 - ex) / (q * q * q) / (sigma * sigma);
      }
    } else {
      const double t = (mu - w[i]) / sigma;
      const double e = exp(-t * t) / sqrt(M_PI);
      s1 += e / sigma;
      s2 += -2.0 * t * e / (sigma * sigma);
    }
  }
  d1 = wave_sum(s1);
  d2 = wave_sum(s2);
}

__global__ void k_occupations(const double* __restrict__ W, int nmo, int nelec, double sigma,
                              int method, int search, int max_cycle, double tol, double mu_shift,
                              double* __restrict__ occ, double* __restrict__ mu_out) {
  const int b = blockIdx.x;
  const int lane = threadIdx.x & 63;
  const double* w = W + (long)b * nmo;
  double mu;
  if (search == 0) {
    double mn = 1e308, mx = -1e308;
    for (int i = lane; i < nmo; i += 64) {
      mn = fmin(mn, w[i]);
      mx = fmax(mx, w[i]);
    }
    for (int off = 32; off >= 1; off >>= 1) {
      mn = fmin(mn, __shfl_xor(mn, off, 64));
      mx = fmax(mx, __shfl_xor(mx, off, 64));
    }
    double lo = mn + mu_shift, hi = mx + mu_shift;
    double ylo = -1.0, yhi = 1.0;
    int loop = 0;
    while (ylo < -tol && yhi > tol && loop < max_cycle) {
      const double mid = (lo + hi) / 2.0;
      const double ym = occ_N(w, nmo, mid, sigma, method, nelec);
      // smearing.py:88-89, written with the same 0/1 weights
      const double nlo = (double)(ym < 0.0) * mid + (double)(ym >= 0.0) * lo;
      const double nhi = (double)(ym >= 0.0) * mid + (double)(ym < 0.0) * hi;
      lo = nlo;
      hi = nhi;
      ylo = occ_N(w, nmo, lo, sigma, method, nelec);
      yhi = occ_N(w, nmo, hi, sigma, method, nelec);
      ++loop;
    }
    // smearing.py:102 return rule, kept verbatim
    mu = (double)(lo > -tol) * lo + (double)(hi < tol) * hi +
         (lo + hi) / 2.0 * (double)(loop == max_cycle);
  } else {
    double mu_old = 1.0;
    mu = w[nelec / 2];
    int loop = 0;
    while (fabs(mu - mu_old) > tol && loop < max_cycle) {
      const double y = occ_N(w, nmo, mu, sigma, method, nelec);
      double d1, d2;
      occ_dN(w, nmo, mu, sigma, method, d1, d2);
      const double g1 = 2.0 * y * d1;
      const double g2 = 2.0 * (d1 * d1 + y * d2);
      mu_old = mu;
      mu = mu - g1 / fabs(g2);
      ++loop;
    }
  }
  for (int i = lane; i < nmo; i += 64) occ[(long)b * nmo + i] = occ_f(mu, w[i], sigma, method);
  if (lane == 0 && mu_out) mu_out[b] = mu;
}

extern "C" int hqc_occupations(const double* w, int B, int nmo, int nelec, double sigma,
                               int method, int search, int max_cycle, double tol, double mu_shift,
                               double* occ, double* mu, void* stream) {
  if (B < 0 || nmo <= 0 || nelec < 0 || !(sigma > 0.0) || method < 0 || method > 1 ||
      search < 0 || search > 1 || max_cycle < 0)
    return HQC_E_ARG;
  if (search == 1 && nelec / 2 >= nmo) return HQC_E_ARG;
  if (B == 0) return HQC_OK;
  hipLaunchKernelGGL(k_occupations, dim3(B), dim3(64), 0, (hipStream_t)stream, w, nmo, nelec,
                     sigma, method, search, max_cycle, tol, mu_shift, occ, mu);
  HQC_CHECK_LAUNCH();
  return HQC_OK;
}

// ---------------------------------------------------------------------------
// Ewald pair potential psi(r) and its gradient (hqc/pbc/potential.py:35-67).
// One block per (pair, geometry); the G sum is a block reduction.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_ewald_pairs(const double* __restrict__ rij, int npair,
                                                     const double* __restrict__ Gv,
                                                     const double* __restrict__ gk, int nG,
                                                     double kappa, double g0,
                                                     double* __restrict__ out) {
  __shared__ double red[4];
  const int p = blockIdx.x, b = blockIdx.y;
  const double* r = rij + ((long)b * npair + p) * 3;
  const double rx = r[0], ry = r[1], rz = r[2];
  double sc = 0.0, gx = 0.0, gy = 0.0, gz = 0.0;
  for (int i = threadIdx.x; i < nG; i += blockDim.x) {
    const double G0 = Gv[3 * i], G1 = Gv[3 * i + 1], G2 = Gv[3 * i + 2];
    double s, c;
    sincos(2.0 * M_PI * (G0 * rx + G1 * ry + G2 * rz), &s, &c);
    const double w = gk[i];
    sc += w * c;
    const double t = -2.0 * M_PI * w * s;
    gx += t * G0;
    gy += t * G1;
    gz += t * G2;
  }
  sc = block_sum(sc, red);
  gx = block_sum(gx, red);
  gy = block_sum(gy, red);
  gz = block_sum(gz, red);
  if (threadIdx.x == 0) {
    const double d = sqrt(rx * rx + ry * ry + rz * rz);
    const double er = erfc(kappa * d);
    const double dsr = -(er / (d * d) + 2.0 * kappa / sqrt(M_PI) * exp(-kappa * kappa * d * d) / d);
    double* o = out + ((long)b * npair + p) * 4;
    o[0] = er / d + sc + g0;
    o[1] = dsr * rx / d + gx;
    o[2] = dsr * ry / d + gy;
    o[3] = dsr * rz / d + gz;
  }
}

extern "C" int hqc_ewald_pairs(const double* rij, int B, int npair, const double* Gv,
                               const double* gk, int nG, double kappa, double g0, double* out,
                               void* stream) {
  if (B < 0 || npair < 0 || nG <= 0) return HQC_E_ARG;
  if (B == 0 || npair == 0) return HQC_OK;
  hipLaunchKernelGGL(k_ewald_pairs, dim3(npair, B), dim3(256), 0, (hipStream_t)stream, rij, npair,
                     Gv, gk, nG, kappa, g0, out);
  HQC_CHECK_LAUNCH();
  return HQC_OK;
}

// Reciprocal part of the same Ewald sum through structure factors: with
// theta_j = 2 pi G.r_j and S(G) = sum_j e^{i theta_j},
//   sum_{i<j} sum_G w_G cos(2 pi G.r_ij)  =  1/2 sum_G w_G (|S(G)|^2 - n),
//   d/dr_i of it                          = -2 pi sum_G w_G G Im(e^{i theta_i} S(G)^*),
// n N_G phase evaluations per geometry instead of n(n-1)/2 N_G.
__global__ __launch_bounds__(256) void k_ewald_sf(const double* __restrict__ frac, int n,
                                                  const double* __restrict__ Gv, int nG,
                                                  double* __restrict__ S) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x, b = blockIdx.y;
  if (k >= nG) return;
  const double G0 = Gv[3 * k], G1 = Gv[3 * k + 1], G2 = Gv[3 * k + 2];
  const double* r = frac + (long)b * n * 3;
  double sc = 0.0, ss = 0.0;
  for (int j = 0; j < n; ++j) {
    double sn, cs;
    sincos(2.0 * M_PI * (G0 * r[3 * j] + G1 * r[3 * j + 1] + G2 * r[3 * j + 2]), &sn, &cs);
    sc += cs;
    ss += sn;
  }
  S[((long)b * nG + k) * 2] = sc;
  S[((long)b * nG + k) * 2 + 1] = ss;
}

// one workgroup per (atom i, geometry): the gradient of atom i, and for i = 0
// the energy sum; fixed-order block reductions
__global__ __launch_bounds__(256) void k_ewald_recip(const double* __restrict__ frac, int n,
                                                     const double* __restrict__ Gv,
                                                     const double* __restrict__ gk, int nG,
                                                     const double* __restrict__ S,
                                                     double* __restrict__ E,
                                                     double* __restrict__ grad) {
  __shared__ double red[4];
  const int i = blockIdx.x, b = blockIdx.y;
  const double* r = frac + ((long)b * n + i) * 3;
  const double rx = r[0], ry = r[1], rz = r[2];
  const double* Sb = S + (long)b * nG * 2;
  double gx = 0.0, gy = 0.0, gz = 0.0, e = 0.0;
  for (int k = threadIdx.x; k < nG; k += blockDim.x) {
    const double G0 = Gv[3 * k], G1 = Gv[3 * k + 1], G2 = Gv[3 * k + 2];
    const double sc = Sb[2 * k], ss = Sb[2 * k + 1], w = gk[k];
    double sn, cs;
    sincos(2.0 * M_PI * (G0 * rx + G1 * ry + G2 * rz), &sn, &cs);
    const double t = -2.0 * M_PI * w * (sn * sc - cs * ss);
    gx += t * G0;
    gy += t * G1;
    gz += t * G2;
    if (i == 0) e += w * (sc * sc + ss * ss - (double)n);
  }
  gx = block_sum(gx, red);
  gy = block_sum(gy, red);
  gz = block_sum(gz, red);
  if (i == 0) e = block_sum(e, red);
  if (threadIdx.x == 0) {
    double* g = grad + ((long)b * n + i) * 3;
    g[0] = gx;
    g[1] = gy;
    g[2] = gz;
    if (i == 0) E[b] = 0.5 * e;
  }
}

extern "C" int hqc_ewald_recip(const double* frac, int B, int n, const double* Gv, const double* gk,
                               int nG, double* E, double* grad, double* work, void* stream) {
  if (B < 0 || n <= 0 || nG <= 0 || !frac || !Gv || !gk || !E || !grad || !work)
    return HQC_E_ARG;
  if (B == 0) return HQC_OK;
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_ewald_sf, dim3(hqc_div_up(nG, 256), B), dim3(256), 0, s, frac, n, Gv, nG,
                     work);
  HQC_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_ewald_recip, dim3(n, B), dim3(256), 0, s, frac, n, Gv, gk, nG, work, E,
                     grad);
  HQC_CHECK_LAUNCH();
  return HQC_OK;
}

extern "C" int hqc_version(void) { return 100; }

// ---------------------------------------------------------------------------
// Row swaps over many tensors in one launch: rows a[k] and b[k] (k < npairs) of
// every tensor exchange places, a tensor being ntensors entries of (base, bytes
// per leading-axis row).  The batched SCF moves the geometries still iterating
// to a prefix of every per-geometry tensor this way: one launch instead of a
// gather and a scatter per tensor.  The pairs must be disjoint (each row in at
// most one pair).  grid (npairs, ntensors), 256 threads per row.
// ---------------------------------------------------------------------------
struct SwapTable {
  char* base[HQC_SWAP_MAX_TENSORS];
  long row_bytes[HQC_SWAP_MAX_TENSORS];
};

__global__ __launch_bounds__(256) void k_swap_rows(SwapTable tb, const long* __restrict__ a,
                                                   const long* __restrict__ b) {
  const int t = blockIdx.y;
  const long rb = tb.row_bytes[t];
  char* ra = tb.base[t] + a[blockIdx.x] * rb;
  char* rbp = tb.base[t] + b[blockIdx.x] * rb;
  if ((((uintptr_t)tb.base[t] | (uintptr_t)rb) & 15) == 0) {
    uint4* x = reinterpret_cast<uint4*>(ra);
    uint4* y = reinterpret_cast<uint4*>(rbp);
    for (long i = threadIdx.x; i < rb / 16; i += blockDim.x) {
      const uint4 u = x[i];
      x[i] = y[i];
      y[i] = u;
    }
  } else if ((((uintptr_t)tb.base[t] | (uintptr_t)rb) & 3) == 0) {
    unsigned* x = reinterpret_cast<unsigned*>(ra);
    unsigned* y = reinterpret_cast<unsigned*>(rbp);
    for (long i = threadIdx.x; i < rb / 4; i += blockDim.x) {
      const unsigned u = x[i];
      x[i] = y[i];
      y[i] = u;
    }
  } else {
    for (long i = threadIdx.x; i < rb; i += blockDim.x) {
      const char u = ra[i];
      ra[i] = rbp[i];
      rbp[i] = u;
    }
  }
}

extern "C" int hqc_swap_rows(void* const* bases, const long* row_bytes, int ntensors,
                             const long* a, const long* b, int npairs, void* stream) {
  if (ntensors < 0 || npairs < 0 || (ntensors && (!bases || !row_bytes))) return HQC_E_ARG;
  if (ntensors > HQC_SWAP_MAX_TENSORS) return HQC_E_SIZE;
  if (ntensors == 0 || npairs == 0) return HQC_OK;
  if (!a || !b || npairs > 65535) return HQC_E_ARG;
  SwapTable tb;
  for (int t = 0; t < ntensors; ++t) {
    if (!bases[t] || row_bytes[t] < 0) return HQC_E_ARG;
    tb.base[t] = static_cast<char*>(bases[t]);
    tb.row_bytes[t] = row_bytes[t];
  }
  hipLaunchKernelGGL(k_swap_rows, dim3(npairs, ntensors), dim3(256), 0, (hipStream_t)stream, tb,
                     a, b);
  HQC_CHECK_LAUNCH();
  return HQC_OK;
}
