// SCF driver behind the C ABI (hqc_scf_create / hqc_scf_run / hqc_scf_destroy):
// the batched Gamma-point restricted Kohn-Sham SCF of make_solver(...,
// dft=True) (hqc/pbc/solver.py:705-767 with the DIIS loop of
// hqc/pbc/scf.py:93-242), driven from C++ on this library's kernels and
// rocBLAS, so that a host without Python can run it.  It follows the engine of
// the Python package step for step (hqc_amd/pbc/engine.py, FP64 path):
//   set-up     S, T by lattice sums, AO values on the grid, v_loc by FFT and
//              V_ep = dV phi diag(v_loc) phi^T, H = T + V_ep, the orthogonaliser
//              v = U s^-1/2 from the eigenpairs of S;
//   Roothaan   F' = v^T F v, Jacobi eigenpairs warm-started in the previous
//              eigenbasis, mo = v C, occupations (aufbau or smearing with the
//              reference's mu search), D = mo f mo^T;
//   Fock       rho in orbital form, v_H by paired FFT-Poisson, LDA v_xc, V =
//              dV phi diag(v_H + v_xc) phi^T, E = Tr(H D) + E_H + E_xc;
//   DIIS       a ring of diis_space (F, SDF - FDS) pairs, diis_start_cycle +
//              diis_space plain iterations first, then the bordered Pulay system
//              (solved on the host per geometry, partial pivoting) for every
//              geometry still running; converged geometries leave the active
//              prefix (their state stays in place), the AO values are read
//              through the geometry index list.
// Outputs are make_solver's 9-tuple in input order (Rydberg where the reference
// reports Rydberg).  Scope: FP64, the Gamma point, LDA functionals, n_ao <= 110
// (the Jacobi kernel), one cell length per context, DIIS on.
#include <rocblas/rocblas.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "common.h"

#define SCF_EIG_TOL_MIN 1e-14
#define SCF_EIG_TOL_MAX 1e-8
#define SCF_EIG_TOL_PER_RES 1e-4
#define SCF_OCC_CUT 1e-15
#define SCF_RY 2.0
#define SCF_EIG_SWEEPS 30

struct hqc_scf_ctx {
  hqc_scf_cell c;
  int nao, nao_pad, N, ld, nT, tmax_img, s_only;
  double dV, mu_shift;
  double* d_alpha;
  double* d_coef;
  int* d_comp;
  double* d_lattice;
  double* d_VG;
  rocblas_handle blas;
};

static long round_up(long x, long m) { return (x + m - 1) / m * m; }

extern "C" int hqc_scf_grid(double L, double grid_length) {
  // Python's round: nearest, ties to even (the default rounding mode)
  return (int)std::nearbyint(L / grid_length / 2.0) * 2 + 1;
}

// ---------------------------------------------------------------------------
// element kernels (grid-stride over `total` entries)
// ---------------------------------------------------------------------------
#define SCF_LOOP(total) \
  for (long idx = (long)blockIdx.x * blockDim.x + threadIdx.x; idx < (total); \
       idx += (long)gridDim.x * blockDim.x)

// out = base + Vpad[:n, :n] (base == NULL: out = Vpad[:n, :n])
__global__ void k_scf_crop_add(const double* __restrict__ Vpad, const double* __restrict__ base,
                               double* __restrict__ out, int n, int npad, long total) {
  const long nn = (long)n * n;
  SCF_LOOP(total) {
    const long b = idx / nn, r = idx % nn;
    const int i = (int)(r / n), j = (int)(r % n);
    const double v = Vpad[(b * npad + i) * npad + j];
    out[idx] = base ? base[idx] + v : v;
  }
}

// out[b,i,j] = U[b,i,j] * s(w[b,j]), s = rsqrt (mode 0) or identity (mode 1)
__global__ void k_scf_scale_cols(const double* __restrict__ U, const double* __restrict__ w,
                                 double* __restrict__ out, int n, long total, int mode) {
  const long nn = (long)n * n;
  SCF_LOOP(total) {
    const long b = idx / nn;
    const int j = (int)(idx % n);
    const double s = w[b * n + j];
    out[idx] = U[idx] * (mode == 0 ? rsqrt(s) : s);
  }
}

// At[b,i,k] = mo[b,i,k] sqrt(max(occ[b,k], 0)) for i < n, k < keep, else 0
__global__ void k_scf_at(const double* __restrict__ mo, const double* __restrict__ occ,
                         double* __restrict__ At, int n, int npad, int keep, int m_pad,
                         long total) {
  SCF_LOOP(total) {
    const int k = (int)(idx % m_pad);
    const long t = idx / m_pad;
    const int i = (int)(t % npad);
    const long b = t / npad;
    At[idx] = (i < n && k < keep) ? mo[(b * n + i) * n + k] * sqrt(fmax(occ[b * n + k], 0.0))
                                  : 0.0;
  }
}

// occ[b,k] = 2 for k < nh, else 0 (hqc/tools/smearing.py:171)
__global__ void k_scf_aufbau(double* __restrict__ occ, int n, int nh, long total) {
  SCF_LOOP(total) occ[idx] = (int)(idx % n) < nh ? 2.0 : 0.0;
}

// e = X - Y
__global__ void k_scf_sub(const double* __restrict__ X, const double* __restrict__ Y,
                          double* __restrict__ e, long total) {
  SCF_LOOP(total) e[idx] = X[idx] - Y[idx];
}

// dst[(b m + s) nn + r] = src[b nn + r]
__global__ void k_scf_put_slot(const double* __restrict__ src, double* __restrict__ dst, int m,
                               int s, long nn, long total) {
  SCF_LOOP(total) {
    const long b = idx / nn, r = idx % nn;
    dst[(b * m + s) * nn + r] = src[idx];
  }
}

// Fg = sum_k ck[b,k] Fk[b,k] (k ascending), then (1 - damp) Fg + damp Fk[b,last]
__global__ void k_scf_extrap(const double* __restrict__ ck, const double* __restrict__ Fk,
                             double* __restrict__ Fg, int m, long nn, int last, double damp,
                             long total) {
  SCF_LOOP(total) {
    const long b = idx / nn, r = idx % nn;
    const double* f = Fk + b * m * nn + r;
    double acc = 0.0;
    for (int k = 0; k < m; ++k) acc += ck[b * m + k] * f[k * nn];
    if (damp != 0.0) acc = (1.0 - damp) * acc + damp * f[last * nn];
    Fg[idx] = acc;
  }
}

// per-geometry reductions, one 256-thread block per geometry (grid.y = slot for
// the DIIS dot products)
__global__ __launch_bounds__(256) void k_scf_trace(const double* __restrict__ A,
                                                   const double* __restrict__ Bm, int n,
                                                   double* __restrict__ out) {
  __shared__ double red[4];
  const long nn = (long)n * n;
  const long b = blockIdx.x;
  double s = 0.0;
  for (long r = threadIdx.x; r < nn; r += 256) {
    const int i = (int)(r / n), j = (int)(r % n);
    s += A[b * nn + r] * Bm[b * nn + (long)j * n + i];
  }
  s = block_sum(s, red);
  if (threadIdx.x == 0) out[b] = s;
}

// d[b, i] = <ek[b, i], e[b]> for slot i = blockIdx.y
__global__ __launch_bounds__(256) void k_scf_dots(const double* __restrict__ ek,
                                                  const double* __restrict__ e, int m, long nn,
                                                  double* __restrict__ d) {
  __shared__ double red[4];
  const long b = blockIdx.x;
  const int i = blockIdx.y;
  const double* x = ek + (b * m + i) * nn;
  const double* y = e + b * nn;
  double s = 0.0;
  for (long r = threadIdx.x; r < nn; r += 256) s += x[r] * y[r];
  s = block_sum(s, red);
  if (threadIdx.x == 0) d[b * m + i] = s;
}

__global__ void k_scf_bh(double* __restrict__ Bh, const double* __restrict__ d, int m, int s,
                         int na) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= na * m) return;
  const int b = idx / m, i = idx % m;
  Bh[((long)b * m + s) * m + i] = d[idx];
  Bh[((long)b * m + i) * m + s] = d[idx];
}

// |e[b]|_F of slot `s` of the ring
__global__ __launch_bounds__(256) void k_scf_fro(const double* __restrict__ ek, int m, int s,
                                                 long nn, double* __restrict__ out) {
  __shared__ double red[4];
  const long b = blockIdx.x;
  const double* x = ek + (b * m + s) * nn;
  double t = 0.0;
  for (long r = threadIdx.x; r < nn; r += 256) t += x[r] * x[r];
  t = block_sum(t, red);
  if (threadIdx.x == 0) out[b] = sqrt(t);
}

// E2[b] = 0.5 dV sum part[b,:,1] + dV sum part[b,:,0]   (E_H + E_xc)
__global__ void k_scf_e2(const double* __restrict__ part, int nblk, double dV, int B,
                         double* __restrict__ E2) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  double s0 = 0.0, s1 = 0.0;
  for (int k = 0; k < nblk; ++k) {
    s0 += part[((long)b * nblk + k) * 2];
    s1 += part[((long)b * nblk + k) * 2 + 1];
  }
  E2[b] = 0.5 * s1 * dV + s0 * dV;
}

__global__ void k_scf_fill_i(int* __restrict__ x, int v, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = v;
}

__global__ void k_scf_iota(int* __restrict__ x, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = i;
}

// rows of 4-byte words: tmp[k] = base[ib[k]], then base[ia[k]] = tmp[k]
__global__ void k_scf_gather(const unsigned* __restrict__ base, unsigned* __restrict__ tmp,
                             const int* __restrict__ ib, long words, long total) {
  SCF_LOOP(total) {
    const long k = idx / words, w = idx % words;
    tmp[idx] = base[(long)ib[k] * words + w];
  }
}
__global__ void k_scf_scatter(unsigned* __restrict__ base, const unsigned* __restrict__ tmp,
                              const int* __restrict__ ia, long words, long total) {
  SCF_LOOP(total) {
    const long k = idx / words, w = idx % words;
    base[(long)ia[k] * words + w] = tmp[idx];
  }
}

// Dp[b] = D[b] zero-padded to npad x npad
__global__ void k_scf_pad(const double* __restrict__ D, double* __restrict__ Dp, int n, int npad,
                          long total) {
  const long pp = (long)npad * npad;
  SCF_LOOP(total) {
    const long b = idx / pp, r = idx % pp;
    const int i = (int)(r / npad), j = (int)(r % npad);
    Dp[idx] = (i < n && j < n) ? D[(b * n + i) * n + j] : 0.0;
  }
}

__global__ void k_scf_add(const double* __restrict__ a, const double* __restrict__ b,
                          double* __restrict__ out, long total) {
  SCF_LOOP(total) out[idx] = a[idx] + b[idx];
}

// Pulay forces from the S/T centre derivatives (forces.py compute_forces):
// F[b,A,c] = sum_{p in A} 2 sum_{q: atom(q) != A} (W_pq dS_pqc - D_pq dT_pqc)
__global__ void k_scf_pulay(const double* __restrict__ W, const double* __restrict__ D,
                            const double* __restrict__ dS, const double* __restrict__ dT, int n,
                            int nloc, int B, double* __restrict__ F) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= B * n * 3) return;
  const int c = idx % 3, A = (idx / 3) % n, b = idx / (3 * n);
  const int nao = n * nloc;
  const long nn = (long)nao * nao;
  double f = 0.0;
  for (int o = 0; o < nloc; ++o) {
    const int p = A * nloc + o;
    double t = 0.0;
    for (int q = 0; q < nao; ++q) {
      if (q / nloc == A) continue;
      const long pq = b * nn + (long)p * nao + q;
      t += W[pq] * dS[pq * 3 + c] - D[pq] * dT[pq * 3 + c];
    }
    f += 2.0 * t;
  }
  F[idx] = f;
}

// fractional coordinates of xp wrapped into the cell (potential.py:38-40)
__global__ void k_scf_frac(const double* __restrict__ xp, double L, double* __restrict__ fr,
                           long total) {
  SCF_LOOP(total) {
    const double x = xp[idx] - L * floor(xp[idx] / L);
    fr[idx] = x / L;
  }
}

// Ewald proton-proton energy (Ry) and forces (Ha/Bohr) of one geometry per block
// (potential.py:35-67 as hqc_amd/pbc/potential.py ewald_pp, Bohr units): the
// real-space pair terms erfc(kappa d)/d over minimum-image fractional
// separations, the reciprocal sums from hqc_ewald_recip, the Madelung term.
__global__ __launch_bounds__(256) void k_scf_ewald(const double* __restrict__ fr, int n, double L,
                                                   double kappa, const double* __restrict__ Erec,
                                                   const double* __restrict__ grec, double mad,
                                                   double* __restrict__ Epp,
                                                   double* __restrict__ Fpp) {
  __shared__ double red[4];
  const int b = blockIdx.x;
  const double* x = fr + (long)b * n * 3;
  const double L2 = L * L, c2 = 2.0 * kappa / sqrt(M_PI);
  double ps = 0.0;
  for (int k = threadIdx.x; k < n * n; k += 256) {
    const int i = k / n, j = k % n;
    if (j <= i) continue;
    double r[3], d2 = 0.0;
    for (int c = 0; c < 3; ++c) {
      r[c] = x[i * 3 + c] - x[j * 3 + c];
      r[c] -= rint(r[c]);
      d2 += r[c] * r[c];
    }
    const double d = sqrt(d2);
    ps += erfc(kappa * d) / d;
  }
  ps = block_sum(ps, red);
  const int npair = n * (n - 1) / 2;
  if (threadIdx.x == 0)
    Epp[b] = 2.0 * (ps + Erec[b] + npair * (-M_PI / (kappa * kappa))) / L + n * mad / L;
  // forces: atom i sums its pair gradients over j in order
  for (int i = threadIdx.x; i < n; i += 256) {
    double f[3] = {0.0, 0.0, 0.0};
    for (int j = 0; j < n; ++j) {
      if (j == i) continue;
      const int a = i < j ? i : j, bj = i < j ? j : i;
      double r[3], d2 = 0.0;
      for (int c = 0; c < 3; ++c) {
        r[c] = x[a * 3 + c] - x[bj * 3 + c];
        r[c] -= rint(r[c]);
        d2 += r[c] * r[c];
      }
      const double d = sqrt(d2), er = erfc(kappa * d);
      const double dsr = -(er / (d * d) + c2 * exp(-(kappa * d) * (kappa * d)) / d);
      const double s = i < j ? -1.0 : 1.0;          // F_i gets -grad_ij (i < j), +grad_ji
      for (int c = 0; c < 3; ++c) f[c] += s * (dsr / d) * r[c] / L2;
    }
    for (int c = 0; c < 3; ++c) Fpp[((long)b * n + i) * 3 + c] = -grec[((long)b * n + i) * 3 + c] / L2 + f[c];
  }
}

// F = F_pulay + (sum_blocks fp - dE + F_pp)
__global__ void k_scf_ftotal(const double* __restrict__ Fp, const double* __restrict__ fp,
                             const double* __restrict__ dE, const double* __restrict__ Fpp,
                             int n, int nblk, int B, double* __restrict__ F) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= B * n * 3) return;
  const int c = idx % 3;
  const long bA = idx / 3;
  double s = 0.0;
  for (int k = 0; k < nblk; ++k) s += fp[(bA * nblk + k) * 3 + c];
  F[idx] = Fp[idx] + (s - dE[idx] + Fpp[idx]);
}

static dim3 grid_of(long total) {
  long b = (total + 255) / 256;
  return dim3((unsigned)std::min(b, 65536L));
}

// ---------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------
static double erfinv_host(double y) {
  // Newton iterations on erf from a rough start (|y| < 1)
  double x = 0.0;
  if (y != 0.0) {
    const double a = 0.147, l = std::log(1.0 - y * y);
    const double t = 2.0 / (M_PI * a) + l / 2.0;
    x = std::copysign(std::sqrt(std::sqrt(t * t - l / a) - t), y);
  }
  for (int it = 0; it < 8; ++it) {
    const double f = std::erf(x) - y;
    x -= f / (2.0 / std::sqrt(M_PI) * std::exp(-x * x));
  }
  return x;
}

template <typename T>
static int to_device(const T* h, size_t n, T** d) {
  if (hipMalloc((void**)d, std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess) return HQC_E_HIP;
  if (n && hipMemcpy(*d, h, n * sizeof(T), hipMemcpyHostToDevice) != hipSuccess)
    return HQC_E_HIP;
  return HQC_OK;
}

extern "C" void hqc_scf_destroy(hqc_scf_ctx* ctx) {
  if (!ctx) return;
  (void)hipFree(ctx->d_alpha);
  (void)hipFree(ctx->d_coef);
  (void)hipFree(ctx->d_comp);
  (void)hipFree(ctx->d_lattice);
  (void)hipFree(ctx->d_VG);
  if (ctx->blas) rocblas_destroy_handle(ctx->blas);
  delete ctx;
}

extern "C" int hqc_scf_create(const hqc_scf_cell* cell, hqc_scf_ctx** out) {
  if (!cell || !out) return HQC_E_ARG;
  *out = nullptr;
  const hqc_scf_cell& c = *cell;
  if (c.n <= 0 || !(c.L > 0.0) || c.g <= 0 || c.na <= 0 || c.nloc <= 0 || c.nl <= 0 ||
      !c.alpha || !c.coef || !c.comp || !(c.rcut > 0.0))
    return HQC_E_ARG;
  if (c.xc_id < 1 || c.xc_id > 5) return HQC_E_ARG;       // LDA functionals only
  if (c.smearing && c.smearing_method != 0 && c.smearing_method != 1) return HQC_E_ARG;
  if (c.search != 0 && c.search != 1) return HQC_E_ARG;
  if (c.max_cycle < 0 || c.diis_space < 1 || c.diis_start_cycle < 0) return HQC_E_ARG;
  const int nao = c.n * c.nloc;
  if (nao < c.n / 2) return HQC_E_ARG;                    // smearing.py:19
  if (nao > 110) return HQC_E_SIZE;                       // the Jacobi eigensolver
  hqc_scf_ctx* x = new hqc_scf_ctx();
  x->c = c;
  x->c.alpha = nullptr;
  x->c.coef = nullptr;
  x->c.comp = nullptr;
  x->nao = nao;
  x->nao_pad = (int)round_up(nao, 16);
  x->N = c.g * c.g * c.g;
  x->ld = (int)round_up(x->N, 64);
  x->dV = std::pow(c.L / c.g, 3.0);
  x->tmax_img = (int)std::floor(24.0 / c.L);              // gto.py:48
  bool s_only = c.na == 4 && c.nloc <= 2;
  for (int o = 0; o < c.nloc; ++o) s_only = s_only && c.comp[o] < 0;
  x->s_only = s_only;
  // lattice translations within rcut (solver.py:168-172), x slowest
  const int tm = (int)std::floor(c.rcut / c.L);
  std::vector<double> lat;
  for (int i = -tm; i <= tm; ++i)
    for (int j = -tm; j <= tm; ++j)
      for (int k = -tm; k <= tm; ++k) {
        const double a = i * c.L, b = j * c.L, d = k * c.L;
        if (a * a + b * b + d * d <= c.rcut * c.rcut) {
          lat.push_back(a);
          lat.push_back(b);
          lat.push_back(d);
        }
      }
  x->nT = (int)(lat.size() / 3);
  // Coulomb kernel 4 pi / (Omega G^2) in FFT order, 0 at G = 0
  const int g = c.g;
  const double omega = std::pow(c.L, 3.0);
  std::vector<double> G1(g), VG((size_t)x->N);
  for (int k = 0; k < g; ++k) G1[k] = (double)(k <= g / 2 ? k : k - g) * 2.0 * M_PI / c.L;
  for (int a = 0; a < g; ++a)
    for (int b = 0; b < g; ++b)
      for (int d = 0; d < g; ++d) {
        const double G2 = G1[a] * G1[a] + G1[b] * G1[b] + G1[d] * G1[d];
        VG[((size_t)a * g + b) * g + d] = (a || b || d) ? 4.0 * M_PI / omega / G2 : 0.0;
      }
  x->mu_shift = 0.0;
  if (c.smearing) {
    x->mu_shift = c.smearing_method == 0
                      ? -c.sigma * std::log(2.0 * nao / c.n - 1.0)              // smearing.py:66
                      : -c.sigma * erfinv_host(1.0 - (double)c.n / nao);
  }
  int err = HQC_OK;
  if (!err) err = to_device(c.alpha, (size_t)c.na, &x->d_alpha);
  if (!err) err = to_device(c.coef, (size_t)c.na * c.nloc, &x->d_coef);
  if (!err) err = to_device(c.comp, (size_t)c.nloc, &x->d_comp);
  if (!err) err = to_device(lat.data(), lat.size(), &x->d_lattice);
  if (!err) err = to_device(VG.data(), VG.size(), &x->d_VG);
  if (!err && rocblas_create_handle(&x->blas) != rocblas_status_success) err = HQC_E_HIP;
  if (err) {
    hqc_scf_destroy(x);
    return err;
  }
  *out = x;
  return HQC_OK;
}

// ---------------------------------------------------------------------------
// run
// ---------------------------------------------------------------------------
namespace {

struct Buf {
  std::vector<void*> all;
  ~Buf() {
    for (void* p : all) (void)hipFree(p);
  }
  template <typename T>
  T* get(size_t n) {
    void* p = nullptr;
    if (hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess) return nullptr;
    all.push_back(p);
    return (T*)p;
  }
};

struct Run {
  hqc_scf_ctx* x;
  hipStream_t st;
  int B, n, nao, np, N, ld, m, nblk;
  long nn;
  size_t vwork_bytes;
  void* vwork;
  // device state (slot order while the SCF runs)
  const double* xp;
  double *P, *phi, *cb, *vl, *Vpad, *S, *T, *Vep, *H, *ws, *U, *vo;
  double *Fk, *ek, *Bh, *E, *Eold, *mo, *w1, *Cw, *Fp, *A, *t1, *t2, *C, *D, *F, *e, *occ, *mu;
  double *At, *rho, *v, *part, *E2, *ck, *dots, *errn, *Fg, *Ki, *Vx;
  int *gidx, *nit, *order, *ia, *ib;
  unsigned* tmp;
  std::vector<double> h_occ;

  int gemm(bool ta, bool tb, const double* Am, long sa, const double* Bm, long sb, double* Cm,
           int batch) {
    // row-major C = op(A) op(B) as column-major C^T = op(B)^T op(A)^T
    const double one = 1.0, zero = 0.0;
    const rocblas_status s = rocblas_dgemm_strided_batched(
        x->blas, tb ? rocblas_operation_transpose : rocblas_operation_none,
        ta ? rocblas_operation_transpose : rocblas_operation_none, nao, nao, nao, &one, Bm, nao,
        sb, Am, nao, sa, &zero, Cm, nao, nn, batch);
    return s == rocblas_status_success ? HQC_OK : HQC_E_HIP;
  }

  // Roothaan step on geometries [0, na): F (stride sf) -> w1, mo, C (warm start
  // in Cw when warm); engine.py _eigh_orth
  int roothaan(const double* Fin, long sf, int na, bool warm, double tol) {
    int r;
    if ((r = gemm(false, false, Fin, sf, vo, nn, t1, na))) return r;          // F v
    if ((r = gemm(true, false, vo, nn, t1, nn, Fp, na))) return r;            // v^T F v
    if (warm) {
      if ((r = gemm(false, false, Fp, nn, Cw, nn, t1, na))) return r;         // F' Cw
      if ((r = gemm(true, false, Cw, nn, t1, nn, A, na))) return r;           // Cw^T F' Cw
      if ((r = hqc_eigh_jacobi(A, na, nao, nullptr, SCF_EIG_SWEEPS, tol, w1, t2, nullptr, st)))
        return r;
      if ((r = gemm(false, false, Cw, nn, t2, nn, C, na))) return r;          // Cw U
    } else {
      if ((r = hqc_eigh_jacobi(Fp, na, nao, nullptr, SCF_EIG_SWEEPS, tol, w1, C, nullptr, st)))
        return r;
    }
    if ((r = gemm(false, false, vo, nn, C, nn, mo, na))) return r;            // v C
    if (hipMemcpyAsync(Cw, C, (size_t)na * nn * sizeof(double), hipMemcpyDeviceToDevice, st) !=
        hipSuccess)
      return HQC_E_HIP;
    return HQC_OK;
  }

  // occ from w1 and D = mo f mo^T on [0, na)
  int density_matrix(int na) {
    const hqc_scf_cell& c = x->c;
    const long tot = (long)na * nao;
    if (c.smearing) {
      int r = hqc_occupations(w1, na, nao, n, c.sigma, c.smearing_method, c.search,
                              c.search_cycle, c.search_tol, x->mu_shift, occ, mu, st);
      if (r) return r;
    } else {
      hipLaunchKernelGGL(k_scf_aufbau, grid_of(tot), dim3(256), 0, st, occ, nao, n / 2, tot);
      HQC_CHECK_LAUNCH();
    }
    hipLaunchKernelGGL(k_scf_scale_cols, grid_of(na * nn), dim3(256), 0, st, mo, occ, t1, nao,
                       (long)na * nn, 1);
    HQC_CHECK_LAUNCH();
    return gemm(false, true, t1, nn, mo, nn, D, na);                           // (mo f) mo^T
  }

  // E2 = E_H + E_xc of the current (mo, occ) on [0, na); V_2 into Vx when wantv
  int fock2e(int na, bool wantv) {
    const hqc_scf_cell& c = x->c;
    // orbitals with f > 1e-15 (their largest count over the batch)
    h_occ.resize((size_t)na * nao);
    if (hipMemcpyAsync(h_occ.data(), occ, h_occ.size() * sizeof(double), hipMemcpyDeviceToHost,
                       st) != hipSuccess || hipStreamSynchronize(st) != hipSuccess)
      return HQC_E_HIP;
    int keep = 1;
    for (int b = 0; b < na; ++b) {
      int k = 0;
      for (int i = 0; i < nao; ++i) k += h_occ[(size_t)b * nao + i] > SCF_OCC_CUT;
      keep = std::max(keep, k);
    }
    const int m_pad = (int)round_up(keep, 16);
    const long ta = (long)na * np * m_pad;
    hipLaunchKernelGGL(k_scf_at, grid_of(ta), dim3(256), 0, st, mo, occ, At, nao, np, keep,
                       m_pad, ta);
    HQC_CHECK_LAUNCH();
    int r = hqc_orb_density_f64(phi, gidx, At, na, np, m_pad, ld, rho, st);
    if (r) return r;
    // v_H: two geometries per complex grid when there are two or more
    const bool paired = na > 1;
    const int ng = paired ? (na + 1) / 2 : na;
    r = paired ? hqc_real_pair_to_complex(rho, na, c.g, ld, x->dV, cb, st)
               : hqc_real_to_complex(rho, na, c.g, ld, x->dV, cb, st);
    if (r || (r = hqc_fft3d_c2c_f64(cb, ng, c.g, -1, st)) ||
        (r = hqc_poisson_mul(cb, x->d_VG, ng, c.g, 0, nullptr, st)) ||
        (r = hqc_fft3d_c2c_f64(cb, ng, c.g, +1, st)) ||
        (r = hqc_xc_eval(rho, cb, paired ? 1 : 0, na, c.g, ld, c.xc_id, v, part, st)))
      return r;
    hipLaunchKernelGGL(k_scf_e2, dim3((na + 255) / 256), dim3(256), 0, st, part,
                       hqc_xc_blocks(na, c.g), x->dV, na,
                       E2);
    HQC_CHECK_LAUNCH();
    if (!wantv) return HQC_OK;
    if ((r = hqc_ao_vmat_f64(phi, gidx, v, na, np, ld, x->dV, Vpad, vwork, vwork_bytes, st)))
      return r;
    hipLaunchKernelGGL(k_scf_crop_add, grid_of(na * nn), dim3(256), 0, st, Vpad, H, F, nao, np,
                       (long)na * nn);                                         // F = H + V_2
    HQC_CHECK_LAUNCH();
    return HQC_OK;
  }

  // one SCF step from Fin: eigenpairs, D, F and SDF - FDS on [0, na); Enew -> E2
  int step(const double* Fin, long sf, int na, bool warm, double tol) {
    int r;
    if ((r = roothaan(Fin, sf, na, warm, tol)) || (r = density_matrix(na)) ||
        (r = fock2e(na, true)))
      return r;
    hipLaunchKernelGGL(k_scf_trace, dim3(na), dim3(256), 0, st, H, D, nao, Ki);  // Tr(H D)
    HQC_CHECK_LAUNCH();
    if ((r = gemm(false, false, S, nn, D, nn, t1, na)) ||                     // S D
        (r = gemm(false, false, t1, nn, F, nn, t2, na)) ||                    // S D F
        (r = gemm(false, false, F, nn, D, nn, t1, na)) ||                     // F D
        (r = gemm(false, false, t1, nn, S, nn, A, na)))                       // F D S
      return r;
    hipLaunchKernelGGL(k_scf_sub, grid_of(na * nn), dim3(256), 0, st, t2, A, e, (long)na * nn);
    HQC_CHECK_LAUNCH();
    return HQC_OK;
  }

  int push(int na, int s) {
    hipLaunchKernelGGL(k_scf_put_slot, grid_of(na * nn), dim3(256), 0, st, F, Fk, m, s, nn,
                       (long)na * nn);
    hipLaunchKernelGGL(k_scf_put_slot, grid_of(na * nn), dim3(256), 0, st, e, ek, m, s, nn,
                       (long)na * nn);
    hipLaunchKernelGGL(k_scf_dots, dim3(na, m), dim3(256), 0, st, ek, e, m, nn, dots);
    hipLaunchKernelGGL(k_scf_bh, dim3((na * m + 255) / 256), dim3(256), 0, st, Bh, dots, m, s,
                       na);
    HQC_CHECK_LAUNCH();
    return HQC_OK;
  }

  // E_old = E; E = Tr(H D) + E2; mo/w1 kept in place (they are the step's)
  int energies(int na) {
    std::vector<double> a((size_t)na), b((size_t)na);
    if (hipMemcpyAsync(Eold, E, na * sizeof(double), hipMemcpyDeviceToDevice, st) != hipSuccess)
      return HQC_E_HIP;
    // E = Ki + E2, formed on the host side of the copy-back below
    if (hipMemcpyAsync(a.data(), Ki, na * sizeof(double), hipMemcpyDeviceToHost, st) !=
            hipSuccess ||
        hipMemcpyAsync(b.data(), E2, na * sizeof(double), hipMemcpyDeviceToHost, st) !=
            hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
      return HQC_E_HIP;
    for (int i = 0; i < na; ++i) a[i] += b[i];
    if (hipMemcpyAsync(E, a.data(), na * sizeof(double), hipMemcpyHostToDevice, st) !=
            hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
      return HQC_E_HIP;
    return HQC_OK;
  }

  // move the kept geometries of [0, na) to a prefix (engine.py _compact)
  int compact(const std::vector<char>& keep, int na, int* new_na) {
    int nk = 0;
    for (int i = 0; i < na; ++i) nk += keep[i] ? 1 : 0;
    *new_na = nk;
    std::vector<int> holes, movers;
    for (int i = 0; i < nk; ++i)
      if (!keep[i]) holes.push_back(i);
    for (int i = nk; i < na; ++i)
      if (keep[i]) movers.push_back(i);
    if (holes.empty()) return HQC_OK;
    std::vector<int> va(holes), vb(movers);
    va.insert(va.end(), movers.begin(), movers.end());
    vb.insert(vb.end(), holes.begin(), holes.end());
    const int cnt = (int)va.size();
    if (hipMemcpyAsync(ia, va.data(), cnt * sizeof(int), hipMemcpyHostToDevice, st) !=
            hipSuccess ||
        hipMemcpyAsync(ib, vb.data(), cnt * sizeof(int), hipMemcpyHostToDevice, st) != hipSuccess)
      return HQC_E_HIP;
    struct Row {
      void* p;
      long bytes;
    };
    const Row rows[] = {{S, nn * 8},  {T, nn * 8},   {Vep, nn * 8},  {H, nn * 8},
                        {vo, nn * 8}, {Fk, m * nn * 8}, {ek, m * nn * 8}, {Bh, (long)m * m * 8},
                        {E, 8},       {Eold, 8},     {mo, nn * 8},   {w1, (long)nao * 8},
                        {Cw, nn * 8}, {gidx, 4},     {nit, 4},       {order, 4}};
    for (const Row& rw : rows) {
      const long words = rw.bytes / 4, tot = words * cnt;
      hipLaunchKernelGGL(k_scf_gather, grid_of(tot), dim3(256), 0, st, (const unsigned*)rw.p, tmp,
                         ib, words, tot);
      hipLaunchKernelGGL(k_scf_scatter, grid_of(tot), dim3(256), 0, st, (unsigned*)rw.p, tmp, ia,
                         words, tot);
    }
    HQC_CHECK_LAUNCH();
    return HQC_OK;
  }
};

// host solve of the bordered Pulay system (m+1 unknowns), LU with partial pivoting
static void pulay_solve(const double* Bm, int m, std::vector<double>& ck) {
  const int M = m + 1;
  std::vector<double> h((size_t)M * M, 0.0), rhs((size_t)M, 0.0);
  double tr = 0.0;
  for (int i = 0; i < m; ++i) tr += Bm[i * m + i];
  const double eta = 1e-12 * (tr / m + 1.0);
  for (int i = 0; i < m; ++i) {
    for (int j = 0; j < m; ++j) h[i * M + j] = Bm[i * m + j] + (i == j ? eta : 0.0);
    h[i * M + m] = -1.0;
    h[m * M + i] = -1.0;
  }
  rhs[m] = -1.0;
  for (int k = 0; k < M; ++k) {
    int p = k;
    for (int i = k + 1; i < M; ++i)
      if (std::fabs(h[i * M + k]) > std::fabs(h[p * M + k])) p = i;
    if (p != k) {
      for (int j = 0; j < M; ++j) std::swap(h[k * M + j], h[p * M + j]);
      std::swap(rhs[k], rhs[p]);
    }
    const double piv = h[k * M + k];
    for (int i = k + 1; i < M; ++i) {
      const double f = piv != 0.0 ? h[i * M + k] / piv : 0.0;
      for (int j = k; j < M; ++j) h[i * M + j] -= f * h[k * M + j];
      rhs[i] -= f * rhs[k];
    }
  }
  std::vector<double> sol((size_t)M, 0.0);
  for (int k = M - 1; k >= 0; --k) {
    double s = rhs[k];
    for (int j = k + 1; j < M; ++j) s -= h[k * M + j] * sol[j];
    sol[k] = h[k * M + k] != 0.0 ? s / h[k * M + k] : 0.0;
  }
  ck.assign(sol.begin(), sol.begin() + m);
}

}  // namespace

namespace {
struct ForceOut {
  double kappa;
  int Gmax;
  double* Etot;
  double* F;
};
}  // namespace

static int scf_forces(Run& R, Buf& buf, const double* xp, const double* Ein_ry,
                      const ForceOut& fo);

static int scf_core(hqc_scf_ctx* x, const double* xp, int B, double* mo_out, double* dm_out,
                    double* bands, double* E_out, double* Ki_out, double* Vep_out,
                    double* Vee_out, double* Se_out, int* conv_out, int* nit_out, void* stream,
                    const ForceOut* fo) {
  if (!x || B < 0 || !xp) return HQC_E_ARG;
  if (B == 0) return HQC_OK;
  const hqc_scf_cell& c = x->c;
  Run R;
  R.x = x;
  R.st = (hipStream_t)stream;
  R.B = B;
  R.n = c.n;
  R.nao = x->nao;
  R.np = x->nao_pad;
  R.N = x->N;
  R.ld = x->ld;
  R.m = c.diis_space;
  R.nn = (long)x->nao * x->nao;
  R.nblk = hqc_xc_blocks(1, c.g);          // the most partial sums per geometry (any na)
  R.xp = xp;
  if (rocblas_set_stream(x->blas, R.st) != rocblas_status_success) return HQC_E_HIP;
  const long nn = R.nn, ld = R.ld, np = R.np, m = R.m;
  Buf buf;
  R.P = buf.get<double>((size_t)B * c.n * 3 * c.g * c.na * c.nl);
  R.phi = buf.get<double>((size_t)B * np * ld);
  R.cb = buf.get<double>((size_t)2 * B * R.N);
  R.vl = buf.get<double>((size_t)B * ld);
  R.Vpad = buf.get<double>((size_t)B * np * np);
  double** mats[] = {&R.S, &R.T, &R.Vep, &R.H, &R.U, &R.vo, &R.mo, &R.Cw, &R.Fp, &R.A,
                     &R.t1, &R.t2, &R.C, &R.D, &R.F, &R.e, &R.Fg, &R.Vx};
  for (double** p : mats) *p = buf.get<double>((size_t)B * nn);
  R.ws = buf.get<double>((size_t)B * R.nao);
  R.w1 = buf.get<double>((size_t)B * R.nao);
  R.occ = buf.get<double>((size_t)B * R.nao);
  R.mu = buf.get<double>((size_t)B);
  R.Fk = buf.get<double>((size_t)B * m * nn);
  R.ek = buf.get<double>((size_t)B * m * nn);
  R.Bh = buf.get<double>((size_t)B * m * m);
  R.E = buf.get<double>((size_t)B);
  R.Eold = buf.get<double>((size_t)B);
  R.At = buf.get<double>((size_t)B * np * np);
  R.rho = buf.get<double>((size_t)B * ld);
  R.v = buf.get<double>((size_t)B * ld);
  R.part = buf.get<double>((size_t)B * R.nblk * 2);
  R.E2 = buf.get<double>((size_t)B);
  R.Ki = buf.get<double>((size_t)B);
  R.ck = buf.get<double>((size_t)B * m);
  R.dots = buf.get<double>((size_t)B * m);
  R.errn = buf.get<double>((size_t)B);
  R.gidx = buf.get<int>((size_t)B);
  R.nit = buf.get<int>((size_t)B);
  R.order = buf.get<int>((size_t)B);
  R.ia = buf.get<int>((size_t)B);
  R.ib = buf.get<int>((size_t)B);
  R.tmp = buf.get<unsigned>((size_t)B * m * nn * 2);
  int r = hqc_ao_vmat_workspace(B, R.np, R.ld, 64, &R.vwork_bytes);
  if (r) return r;
  R.vwork = buf.get<char>(R.vwork_bytes);
  for (void* p : buf.all)
    if (!p) return HQC_E_HIP;
  hipStream_t st = R.st;
  if (hipMemsetAsync(R.phi, 0, (size_t)B * np * ld * sizeof(double), st) != hipSuccess ||
      hipMemsetAsync(R.vl, 0, (size_t)B * ld * sizeof(double), st) != hipSuccess ||
      hipMemsetAsync(R.rho, 0, (size_t)B * ld * sizeof(double), st) != hipSuccess ||
      hipMemsetAsync(R.v, 0, (size_t)B * ld * sizeof(double), st) != hipSuccess)
    return HQC_E_HIP;
  // ---- set-up (engine.py setup)
  r = x->s_only ? hqc_overlap_kinetic_s(xp, B, c.n, x->d_lattice, x->nT, x->d_alpha, c.na,
                                        x->d_coef, c.nloc, R.S, R.T, nullptr, nullptr, st)
                : hqc_overlap_kinetic(xp, B, c.n, x->d_lattice, x->nT, x->d_alpha, c.na,
                                      x->d_coef, x->d_comp, c.nloc, R.S, R.T, nullptr, nullptr,
                                      st);
  if (r || (r = hqc_gto_tables(xp, B, c.n, c.g, c.L, x->tmax_img, x->d_alpha, c.na, c.nl, R.P,
                               st)) ||
      (r = hqc_ao_values_f64(R.P, B, c.n, c.g, c.na, c.nl, x->d_coef, x->d_comp, c.nloc, R.np,
                             R.ld, R.phi, st)) ||
      (r = hqc_vloc_G(xp, B, c.n, c.g, c.L, x->d_VG, R.cb, st)) ||
      (r = hqc_fft3d_c2c_f64(R.cb, B, c.g, +1, st)) ||
      (r = hqc_complex_real(R.cb, B, c.g, R.ld, R.vl, st)) ||
      (r = hqc_ao_vmat_f64(R.phi, nullptr, R.vl, B, R.np, R.ld, x->dV, R.Vpad, R.vwork,
                           R.vwork_bytes, st)))
    return r;
  hipLaunchKernelGGL(k_scf_crop_add, grid_of(B * nn), dim3(256), 0, st, R.Vpad, nullptr, R.Vep,
                     R.nao, R.np, (long)B * nn);
  hipLaunchKernelGGL(k_scf_crop_add, grid_of(B * nn), dim3(256), 0, st, R.Vpad, R.T, R.H, R.nao,
                     R.np, (long)B * nn);
  HQC_CHECK_LAUNCH();
  // orthogonaliser v = U s^-1/2 (solver.py:729-730)
  if ((r = hqc_eigh_jacobi(R.S, B, R.nao, nullptr, SCF_EIG_SWEEPS, SCF_EIG_TOL_MIN, R.ws, R.U,
                           nullptr, st)))
    return r;
  hipLaunchKernelGGL(k_scf_scale_cols, grid_of(B * nn), dim3(256), 0, st, R.U, R.ws, R.vo, R.nao,
                     (long)B * nn, 0);
  hipLaunchKernelGGL(k_scf_iota, dim3((B + 255) / 256), dim3(256), 0, st, R.gidx, B);
  hipLaunchKernelGGL(k_scf_iota, dim3((B + 255) / 256), dim3(256), 0, st, R.order, B);
  HQC_CHECK_LAUNCH();
  // ---- SCF (engine.py scf, DIIS branch)
  if ((r = R.roothaan(R.H, nn, B, false, SCF_EIG_TOL_MIN)) || (r = R.density_matrix(B)) ||
      (r = R.fock2e(B, true)))
    return r;
  if ((r = R.gemm(false, false, R.S, nn, R.D, nn, R.t1, B)) ||
      (r = R.gemm(false, false, R.t1, nn, R.F, nn, R.t2, B)) ||
      (r = R.gemm(false, false, R.F, nn, R.D, nn, R.t1, B)) ||
      (r = R.gemm(false, false, R.t1, nn, R.S, nn, R.A, B)))
    return r;
  hipLaunchKernelGGL(k_scf_sub, grid_of(B * nn), dim3(256), 0, st, R.t2, R.A, R.e, (long)B * nn);
  for (int s = 0; s < m; ++s) {
    hipLaunchKernelGGL(k_scf_put_slot, grid_of(B * nn), dim3(256), 0, st, R.F, R.Fk, (int)m, s,
                       nn, (long)B * nn);
    hipLaunchKernelGGL(k_scf_put_slot, grid_of(B * nn), dim3(256), 0, st, R.e, R.ek, (int)m, s,
                       nn, (long)B * nn);
  }
  // Bh = <e0, e0> everywhere
  hipLaunchKernelGGL(k_scf_dots, dim3(B, 1), dim3(256), 0, st, R.ek, R.e, (int)m, nn, R.dots);
  HQC_CHECK_LAUNCH();
  {
    std::vector<double> d((size_t)B * m), bh((size_t)B * m * m);
    if (hipMemcpyAsync(d.data(), R.dots, d.size() * sizeof(double), hipMemcpyDeviceToHost, st) !=
            hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
      return HQC_E_HIP;
    for (int b = 0; b < B; ++b)
      for (long k = 0; k < m * m; ++k) bh[(size_t)b * m * m + k] = d[(size_t)b * m];
    if (hipMemcpyAsync(R.Bh, bh.data(), bh.size() * sizeof(double), hipMemcpyHostToDevice, st) !=
            hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
      return HQC_E_HIP;
  }
  {
    std::vector<double> ones((size_t)B, 1.0);
    if (hipMemcpyAsync(R.E, ones.data(), B * sizeof(double), hipMemcpyHostToDevice, st) !=
            hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
      return HQC_E_HIP;
  }
  int last = (int)m - 1;
  int na = B;
  for (int it = 0; it < c.diis_start_cycle + m; ++it) {                      // scf.py:130-172
    if ((r = R.step(R.Fk + (long)last * nn, m * nn, na, true, SCF_EIG_TOL_MAX))) return r;
    last = (last + 1) % (int)m;
    if ((r = R.push(na, last)) || (r = R.energies(na))) return r;
  }
  int loop = c.diis_start_cycle + (int)m;
  hipLaunchKernelGGL(k_scf_fill_i, dim3((B + 255) / 256), dim3(256), 0, st, R.nit, loop, B);
  HQC_CHECK_LAUNCH();
  std::vector<double> hE, hEo, hErr, hBh;
  double res = -1.0;
  auto cont = [&](int na_, std::vector<char>& keep) -> int {
    hE.resize(na_);
    hEo.resize(na_);
    hErr.resize(na_);
    if (na_ == 0) return HQC_OK;
    hipLaunchKernelGGL(k_scf_fro, dim3(na_), dim3(256), 0, st, R.ek, (int)m, last, nn, R.errn);
    HQC_CHECK_LAUNCH();
    if (hipMemcpyAsync(hE.data(), R.E, na_ * sizeof(double), hipMemcpyDeviceToHost, st) !=
            hipSuccess ||
        hipMemcpyAsync(hEo.data(), R.Eold, na_ * sizeof(double), hipMemcpyDeviceToHost, st) !=
            hipSuccess ||
        hipMemcpyAsync(hErr.data(), R.errn, na_ * sizeof(double), hipMemcpyDeviceToHost, st) !=
            hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
      return HQC_E_HIP;
    keep.assign(na_, 0);
    res = hErr[0];
    for (int i = 0; i < na_; ++i) {
      res = std::min(res, hErr[i]);
      const double dE = std::fabs(hE[i] - hEo[i]);
      keep[i] = loop < c.max_cycle && (dE > c.tol || hErr[i] > 10 * c.tol);
    }
    return HQC_OK;
  };
  // E_old = E - 1 after the warm-up (scf.py:172)
  {
    std::vector<double> e((size_t)B);
    if (hipMemcpyAsync(e.data(), R.E, B * sizeof(double), hipMemcpyDeviceToHost, st) !=
            hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
      return HQC_E_HIP;
    for (double& q : e) q -= 1.0;
    if (hipMemcpyAsync(R.Eold, e.data(), B * sizeof(double), hipMemcpyHostToDevice, st) !=
            hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
      return HQC_E_HIP;
  }
  std::vector<char> keep;
  if ((r = cont(na, keep)) || (r = R.compact(keep, na, &na))) return r;
  std::vector<double> ck, hck;
  while (na > 0) {                                                             // scf.py:174-239
    hBh.resize((size_t)na * m * m);
    if (hipMemcpyAsync(hBh.data(), R.Bh, hBh.size() * sizeof(double), hipMemcpyDeviceToHost, st) !=
            hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
      return HQC_E_HIP;
    hck.resize((size_t)na * m);
    for (int b = 0; b < na; ++b) {
      pulay_solve(hBh.data() + (size_t)b * m * m, (int)m, ck);
      std::copy(ck.begin(), ck.end(), hck.begin() + (size_t)b * m);
    }
    if (hipMemcpyAsync(R.ck, hck.data(), hck.size() * sizeof(double), hipMemcpyHostToDevice, st) !=
        hipSuccess)
      return HQC_E_HIP;
    hipLaunchKernelGGL(k_scf_extrap, grid_of(na * nn), dim3(256), 0, st, R.ck, R.Fk, R.Fg, (int)m,
                       nn, last, c.diis_damp, (long)na * nn);
    HQC_CHECK_LAUNCH();
    const double etol =
        std::min(SCF_EIG_TOL_MAX, std::max(SCF_EIG_TOL_MIN, SCF_EIG_TOL_PER_RES * res));
    if ((r = R.step(R.Fg, nn, na, true, etol))) return r;
    last = (last + 1) % (int)m;
    if ((r = R.push(na, last)) || (r = R.energies(na))) return r;
    ++loop;
    hipLaunchKernelGGL(k_scf_fill_i, dim3((na + 255) / 256), dim3(256), 0, st, R.nit, loop, na);
    HQC_CHECK_LAUNCH();
    if ((r = cont(na, keep)) || (r = R.compact(keep, na, &na))) return r;
  }
  // ---- results (engine.py solve): D, occ of the final orbitals, E_2, traces
  if ((r = R.density_matrix(B)) || (r = R.fock2e(B, false))) return r;
  hipLaunchKernelGGL(k_scf_trace, dim3(B), dim3(256), 0, st, R.T, R.D, R.nao, R.Ki);
  hipLaunchKernelGGL(k_scf_trace, dim3(B), dim3(256), 0, st, R.Vep, R.D, R.nao, R.Fg);
  HQC_CHECK_LAUNCH();
  std::vector<double> hKi(B), hVep(B), hE2(B), hEf(B), hw((size_t)B * R.nao),
      hocc((size_t)B * R.nao);
  std::vector<int> hnit(B), hord(B);
  if (hipMemcpyAsync(hKi.data(), R.Ki, B * 8, hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipMemcpyAsync(hVep.data(), R.Fg, B * 8, hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipMemcpyAsync(hE2.data(), R.E2, B * 8, hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipMemcpyAsync(hEf.data(), R.E, B * 8, hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipMemcpyAsync(hw.data(), R.w1, hw.size() * 8, hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipMemcpyAsync(hocc.data(), R.occ, hocc.size() * 8, hipMemcpyDeviceToHost, st) !=
          hipSuccess ||
      hipMemcpyAsync(hnit.data(), R.nit, B * 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipMemcpyAsync(hord.data(), R.order, B * 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return HQC_E_HIP;
  // per-geometry scalars in input order
  std::vector<double> oE(B), oKi(B), oVep(B), oVee(B), oSe(B), ow((size_t)B * R.nao);
  std::vector<int> oconv(B), onit(B);
  for (int i = 0; i < B; ++i) {
    const int g = hord[i];
    oE[g] = hEf[i] * SCF_RY;
    oKi[g] = hKi[i] * SCF_RY;
    oVep[g] = hVep[i] * SCF_RY;
    oVee[g] = hE2[i] * SCF_RY;
    double sa = 0.0, sb = 0.0;                                                 // solver.py:579-590
    for (int k = 0; k < R.nao; ++k) {
      const double h = hocc[(size_t)i * R.nao + k] / 2.0;
      if (h > SCF_OCC_CUT) sa += h * std::log(h);
      if (1.0 - h > SCF_OCC_CUT) sb += (1.0 - h) * std::log(1.0 - h);
      ow[(size_t)g * R.nao + k] = hw[(size_t)i * R.nao + k] * SCF_RY;
    }
    oSe[g] = -2.0 * (sa + sb);
    oconv[g] = hnit[i] != c.max_cycle;
    onit[g] = hnit[i];
  }
  auto put = [&](void* dst, const void* src, size_t bytes) {
    return !dst || hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st) == hipSuccess;
  };
  if (!put(bands, ow.data(), ow.size() * 8) || !put(E_out, oE.data(), B * 8) ||
      !put(Ki_out, oKi.data(), B * 8) || !put(Vep_out, oVep.data(), B * 8) ||
      !put(Vee_out, oVee.data(), B * 8) || !put(Se_out, oSe.data(), B * 8) ||
      !put(conv_out, oconv.data(), B * 4) || !put(nit_out, onit.data(), B * 4))
    return HQC_E_HIP;
  // matrices in input order: row i of the slot arrays goes to geometry order[i]
  const long words = nn * 2;
  auto unperm = [&](const double* src, double* dst) -> int {
    if (!dst) return HQC_OK;
    hipLaunchKernelGGL(k_scf_scatter, grid_of(B * words), dim3(256), 0, st, (unsigned*)dst,
                       (const unsigned*)src, R.order, words, (long)B * words);
    HQC_CHECK_LAUNCH();
    return HQC_OK;
  };
  if ((r = unperm(R.mo, mo_out)) || (r = unperm(R.D, dm_out))) return r;
  if (fo) {
    // the force pass works in input order: mo and w1 of each geometry back there
    if ((r = unperm(R.mo, R.t1))) return r;
    hipLaunchKernelGGL(k_scf_scatter, grid_of((long)B * R.nao * 2), dim3(256), 0, st,
                       (unsigned*)R.ws, (const unsigned*)R.w1, R.order, (long)R.nao * 2,
                       (long)B * R.nao * 2);
    HQC_CHECK_LAUNCH();
    if (hipMemcpyAsync(R.mo, R.t1, (size_t)B * nn * 8, hipMemcpyDeviceToDevice, st) !=
            hipSuccess ||
        hipMemcpyAsync(R.w1, R.ws, (size_t)B * R.nao * 8, hipMemcpyDeviceToDevice, st) !=
            hipSuccess)
      return HQC_E_HIP;
    if ((r = scf_forces(R, buf, xp, oE.data(), *fo))) return r;
  }
  if (hipStreamSynchronize(st) != hipSuccess) return HQC_E_HIP;
  return HQC_OK;
}

// Forces of the converged states in input order (forces.py compute_forces, DFT
// at Gamma): Pulay terms from the S/T centre derivatives with W = mo (f e) mo^T,
// the grid adjoint of the AO contractions (X = D phi, v_eff = v_H + v_xc + v_loc),
// the structure-factor derivative of the local potential, and Ewald.
static int scf_forces(Run& R, Buf& buf, const double* xp, const double* Ein_ry,
                      const ForceOut& fo) {
  hqc_scf_ctx* x = R.x;
  const hqc_scf_cell& c = x->c;
  hipStream_t st = R.st;
  const int B = R.B, n = c.n, nao = R.nao, np = R.np;
  const long nn = R.nn, ld = R.ld;
  const int nlf = std::max(3, c.nl + 1);           // l_max + 2 powers, at least 3
  const int nblkF = hqc_force_blocks(c.g);
  // Ewald G vectors and weights (potential.py:6-33)
  std::vector<double> Gv, gk;
  double mad = 0.0;
  for (int i = -fo.Gmax; i <= fo.Gmax; ++i)
    for (int j = -fo.Gmax; j <= fo.Gmax; ++j)
      for (int k = -fo.Gmax; k <= fo.Gmax; ++k) {
        const int G2 = i * i + j * j + k * k;
        if (G2 == 0 || G2 > fo.Gmax * fo.Gmax) continue;
        Gv.push_back(i);
        Gv.push_back(j);
        Gv.push_back(k);
        const double w =
            std::exp(-M_PI * M_PI * G2 / (fo.kappa * fo.kappa)) / (M_PI * G2);
        gk.push_back(w);
        mad += w;
      }
  mad += -M_PI / (fo.kappa * fo.kappa) - 2.0 * fo.kappa / std::sqrt(M_PI);
  const int nG = (int)gk.size();
  double* X = buf.get<double>((size_t)B * np * ld);
  double* Dp = buf.get<double>((size_t)B * np * np);
  double* dS = buf.get<double>((size_t)B * nn * 3);
  double* dT = buf.get<double>((size_t)B * nn * 3);
  double* P3 = buf.get<double>((size_t)B * n * 3 * c.g * c.na * nlf);
  double* fp = buf.get<double>((size_t)B * n * nblkF * 3);
  double* Fpul = buf.get<double>((size_t)B * n * 3);
  double* dE = buf.get<double>((size_t)B * n * 3);
  double* Fpp = buf.get<double>((size_t)B * n * 3);
  double* grec = buf.get<double>((size_t)B * n * 3);
  double* fr = buf.get<double>((size_t)B * n * 3);
  double* Erec = buf.get<double>((size_t)B);
  double* Epp = buf.get<double>((size_t)B);
  double* work = buf.get<double>((size_t)B * nG * 2);
  double* rg = buf.get<double>((size_t)2 * B * R.N);
  double* dG = nullptr;
  double* dgk = nullptr;
  if (to_device(Gv.data(), Gv.size(), &dG) || to_device(gk.data(), gk.size(), &dgk)) {
    (void)hipFree(dG);
    (void)hipFree(dgk);
    return HQC_E_HIP;
  }
  buf.all.push_back(dG);
  buf.all.push_back(dgk);
  for (void* p : buf.all)
    if (!p) return HQC_E_HIP;
  int r;
  // D and W = mo diag(f e) mo^T of the final orbitals (input order, phi by identity)
  hipLaunchKernelGGL(k_scf_iota, dim3((B + 255) / 256), dim3(256), 0, st, R.gidx, B);
  HQC_CHECK_LAUNCH();
  if ((r = R.density_matrix(B))) return r;
  hipLaunchKernelGGL(k_scf_scale_cols, grid_of(B * nn), dim3(256), 0, st, R.mo, R.occ, R.t1, nao,
                     (long)B * nn, 1);
  hipLaunchKernelGGL(k_scf_scale_cols, grid_of(B * nn), dim3(256), 0, st, R.t1, R.w1, R.t2, nao,
                     (long)B * nn, 1);
  HQC_CHECK_LAUNCH();
  if ((r = R.gemm(false, true, R.t2, nn, R.mo, nn, R.A, B))) return r;       // W
  // S/T centre derivatives
  r = x->s_only ? hqc_overlap_kinetic_s(xp, B, n, x->d_lattice, x->nT, x->d_alpha, c.na,
                                        x->d_coef, c.nloc, R.S, R.T, dS, dT, st)
                : hqc_overlap_kinetic(xp, B, n, x->d_lattice, x->nT, x->d_alpha, c.na,
                                      x->d_coef, x->d_comp, c.nloc, R.S, R.T, dS, dT, st);
  if (r) return r;
  hipLaunchKernelGGL(k_scf_pulay, dim3((B * n * 3 + 255) / 256), dim3(256), 0, st, R.A, R.D, dS,
                     dT, n, c.nloc, B, Fpul);
  hipLaunchKernelGGL(k_scf_pad, grid_of((long)B * np * np), dim3(256), 0, st, R.D, Dp, nao, np,
                     (long)B * np * np);
  HQC_CHECK_LAUNCH();
  // grid adjoint: rho, X = D phi, v_H with rho~(G) kept, v_xc, v_eff
  if ((r = hqc_ao_density_f64(R.phi, nullptr, Dp, B, np, ld, R.rho, X, st)) ||
      (r = hqc_real_to_complex(R.rho, B, c.g, ld, x->dV, R.cb, st)) ||
      (r = hqc_fft3d_c2c_f64(R.cb, B, c.g, -1, st)) ||
      (r = hqc_poisson_mul(R.cb, x->d_VG, B, c.g, 0, rg, st)) ||
      (r = hqc_fft3d_c2c_f64(R.cb, B, c.g, +1, st)) ||
      (r = hqc_xc_eval(R.rho, R.cb, 0, B, c.g, ld, c.xc_id, R.v, R.part, st)))
    return r;
  hipLaunchKernelGGL(k_scf_add, grid_of(B * ld), dim3(256), 0, st, R.v, R.vl, R.v, (long)B * ld);
  HQC_CHECK_LAUNCH();
  if ((r = hqc_gto_tables(xp, B, n, c.g, c.L, x->tmax_img, x->d_alpha, c.na, nlf, P3, st)) ||
      (r = hqc_grid_forces(P3, B, n, c.g, c.na, nlf, x->d_alpha, x->d_coef, x->d_comp, c.nloc,
                           np, (int)ld, X, R.v, nullptr, x->dV, fp, st)) ||
      (r = hqc_vloc_force(xp, B, n, c.g, c.L, x->d_VG, rg, dE, st)))
    return r;
  // Ewald in Bohr units
  hipLaunchKernelGGL(k_scf_frac, grid_of((long)B * n * 3), dim3(256), 0, st, xp, c.L, fr,
                     (long)B * n * 3);
  HQC_CHECK_LAUNCH();
  if ((r = hqc_ewald_recip(fr, B, n, dG, dgk, nG, Erec, grec, work, st))) return r;
  hipLaunchKernelGGL(k_scf_ewald, dim3(B), dim3(256), 0, st, fr, n, c.L, fo.kappa, Erec, grec,
                     mad, Epp, Fpp);
  hipLaunchKernelGGL(k_scf_ftotal, dim3((B * n * 3 + 255) / 256), dim3(256), 0, st, Fpul, fp, dE,
                     Fpp, n, nblkF, B, fo.F);
  HQC_CHECK_LAUNCH();
  // E_tot = E + V_pp (Ry)
  std::vector<double> hEpp((size_t)B);
  if (hipMemcpyAsync(hEpp.data(), Epp, B * 8, hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return HQC_E_HIP;
  for (int i = 0; i < B; ++i) hEpp[i] += Ein_ry[i];
  if (fo.Etot && hipMemcpyAsync(fo.Etot, hEpp.data(), B * 8, hipMemcpyHostToDevice, st) !=
                     hipSuccess)
    return HQC_E_HIP;
  return hipStreamSynchronize(st) == hipSuccess ? HQC_OK : HQC_E_HIP;
}

extern "C" int hqc_scf_run(hqc_scf_ctx* x, const double* xp, int B, double* mo_out,
                           double* dm_out, double* bands, double* E_out, double* Ki_out,
                           double* Vep_out, double* Vee_out, double* Se_out, int* conv_out,
                           int* nit_out, void* stream) {
  return scf_core(x, xp, B, mo_out, dm_out, bands, E_out, Ki_out, Vep_out, Vee_out, Se_out,
                  conv_out, nit_out, stream, nullptr);
}

extern "C" int hqc_scf_forces(hqc_scf_ctx* x, const double* xp, int B, double kappa, int Gmax,
                              double* E_tot, double* F, int* conv_out, int* nit_out,
                              void* stream) {
  if (!x || !F || !(kappa > 0.0) || Gmax < 1) return HQC_E_ARG;
  // the forces are -dA/dR only for Fermi occupations (forces.py check_supported)
  if (x->c.smearing && x->c.smearing_method != 0) return HQC_E_ARG;
  const ForceOut fo{kappa, Gmax, E_tot, F};
  return scf_core(x, xp, B, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                  conv_out, nit_out, stream, &fo);
}
