// Analytic periodic overlap and kinetic matrices (and their first-argument
// derivatives for forces) by lattice sums.
//
// Reference: hqc/pbc/solver.py:233-389.  For primitive exponents a (on atom
// A) and b (on atom B) with mu = ab/(a+b) and R = R_A - R_B + T, the s-s
// overlap is pi^1.5 (a+b)^-1.5 exp(-mu R^2) and the kinetic integral is
// -1/2 Laplacian_R of it (= mu(3 - 2 mu R^2) S, solver.py:252-254).  A p_i
// function on A is (1/2a) d/dA_i of the s function and a p_j on B is
// (1/2b) d/dB_j (solver.py:307-341); since d/dA = d/dR and d/dB = -d/dR,
// every needed integral is a mixed R-derivative of exp(-mu R^2), i.e. a
// product over x,y,z of Hermite factors e_k(R_d):
//   d^k/dR^k exp(-mu R^2) = e_k(R) exp(-mu R^2),
//   e_0 = 1, e_1 = -2 mu R, e_{k+1} = -2 mu R e_k - 2 k mu e_{k-1}.
#include "common.h"

#define INT_MAX_ORDER 5
#define INT_MAX_NA 16
#define INT_MAX_NLOC 16

// Hermite factors up to order nmax (<= INT_MAX_ORDER); higher entries untouched.
__device__ __forceinline__ void hermite(double mu, double R, double* e, int nmax) {
  e[0] = 1.0;
  e[1] = -2.0 * mu * R;
  for (int k = 1; k < nmax; ++k) e[k + 1] = -2.0 * mu * R * e[k] - 2.0 * k * mu * e[k - 1];
}

// Lattice terms with mu |R|^2 beyond this contribute < e^-46 ~ 1e-20 of a
// unit-normalised integral and are skipped.
#define INT_EXP_CUT 46.0

// One thread per (b, A, Bat, o, q) with (A, o) <= (Bat, q); each writes its entry
// and the mirrored one.  CPLX: k-point version, every lattice term
// carries the Bloch phase e^{-i k.T} (hqc/pbc/solver.py:270-271) and the
// outputs are complex (interleaved re, im).
template <bool CPLX>
__global__ __launch_bounds__(256) void k_overlap_kinetic(
    const double* __restrict__ xp, const double* __restrict__ kpt, long total, int n,
    const double* __restrict__ lat, int nT, const double* __restrict__ alpha, int na,
    const double* __restrict__ coef, const int* __restrict__ comp, int nloc,
    double* __restrict__ S, double* __restrict__ Tm, double* __restrict__ dS,
    double* __restrict__ dT) {
  long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= total) return;
  // threads run over the unordered atom pairs A <= Bat only (no lane waits out
  // the lattice loop for a mirrored entry); (Bat,q | A,o) is written as the mirror
  const int q = (int)(idx % nloc);
  long t1 = idx / nloc;
  const int o = (int)(t1 % nloc);
  t1 /= nloc;
  const long npair = (long)n * (n + 1) / 2;
  const int pr = (int)(t1 % npair);
  const int b = (int)(t1 / npair);
  // row A of the upper triangle holds pairs [A n - A(A-1)/2, (A+1) n - A(A+1)/2)
  int A = (int)((2.0 * n + 1.0 - sqrt((2.0 * n + 1.0) * (2.0 * n + 1.0) - 8.0 * pr)) * 0.5);
  if (A < 0) A = 0;
  if (A > n - 1) A = n - 1;
  while (A > 0 && A * n - A * (A - 1) / 2 > pr) --A;
  while (A < n - 1 && (A + 1) * n - (A + 1) * A / 2 <= pr) ++A;
  const int Bat = A + (pr - (A * n - A * (A - 1) / 2));
  if (Bat == A && q < o) return;
  const double* xA = xp + ((long)b * n + A) * 3;
  const double* xB = xp + ((long)b * n + Bat) * 3;
  const double d0 = xA[0] - xB[0], d1 = xA[1] - xB[1], d2 = xA[2] - xB[2];
  double k0 = 0.0, k1 = 0.0, k2 = 0.0;
  if (CPLX) {
    k0 = kpt[b * 3];
    k1 = kpt[b * 3 + 1];
    k2 = kpt[b * 3 + 2];
  }
  const int co = comp[o], cq = comp[q];
  int base[3] = {0, 0, 0};
  if (co >= 0) base[co] += 1;
  if (cq >= 0) base[cq] += 1;
  const bool deriv = dS != nullptr;
  int nmax[3];
  for (int d = 0; d < 3; ++d) nmax[d] = base[d] + 2 + (deriv ? 1 : 0);
  // [0] real part, [1] imaginary part (CPLX only)
  double s[2] = {0.0, 0.0}, t[2] = {0.0, 0.0};
  double ds[2][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}}, dt[2][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};
  for (int ia = 0; ia < na; ++ia) {
    const double wa = coef[ia * nloc + o];
    if (wa == 0.0) continue;
    for (int ib = 0; ib < na; ++ib) {
      const double wb = coef[ib * nloc + q];
      if (wb == 0.0) continue;
      const double a = alpha[ia], bb = alpha[ib];
      const double mu = a * bb / (a + bb);
      double fac = wa * wb * M_PI * sqrt(M_PI) * pow(a + bb, -1.5);
      if (co >= 0) fac *= 1.0 / (2.0 * a);
      if (cq >= 0) fac *= -1.0 / (2.0 * bb);
      double ss[2] = {0.0, 0.0}, tt[2] = {0.0, 0.0};
      double dss[2][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};
      double dtt[2][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};
      for (int it = 0; it < nT; ++it) {
        const double* T3 = lat + 3 * it;
        const double R[3] = {d0 + T3[0], d1 + T3[1], d2 + T3[2]};
        const double mr2 = mu * (R[0] * R[0] + R[1] * R[1] + R[2] * R[2]);
        if (mr2 > INT_EXP_CUT) continue;
        double ph[2] = {1.0, 0.0};
        if (CPLX) {
          double sn, cs;
          sincos(k0 * T3[0] + k1 * T3[1] + k2 * T3[2], &sn, &cs);
          ph[0] = cs;
          ph[1] = -sn;
        }
        const double E = exp(-mr2);
        double h[3][INT_MAX_ORDER + 1];
        hermite(mu, R[0], h[0], nmax[0]);
        hermite(mu, R[1], h[1], nmax[1]);
        hermite(mu, R[2], h[2], nmax[2]);
        const double hx = h[0][base[0]], hy = h[1][base[1]], hz = h[2][base[2]];
        const double vs = E * hx * hy * hz;
        const double vt = E * (h[0][base[0] + 2] * hy * hz + hx * h[1][base[1] + 2] * hz +
                               hx * hy * h[2][base[2] + 2]);
        double vds[3], vdt[3];
        if (deriv) {
          const double gx = h[0][base[0] + 1], gy = h[1][base[1] + 1], gz = h[2][base[2] + 1];
          vds[0] = E * gx * hy * hz;
          vds[1] = E * hx * gy * hz;
          vds[2] = E * hx * hy * gz;
          vdt[0] = E * (h[0][base[0] + 3] * hy * hz + gx * h[1][base[1] + 2] * hz +
                        gx * hy * h[2][base[2] + 2]);
          vdt[1] = E * (h[0][base[0] + 2] * gy * hz + hx * h[1][base[1] + 3] * hz +
                        hx * gy * h[2][base[2] + 2]);
          vdt[2] = E * (h[0][base[0] + 2] * hy * gz + hx * h[1][base[1] + 2] * gz +
                        hx * hy * h[2][base[2] + 3]);
        }
#pragma unroll
        for (int part = 0; part < (CPLX ? 2 : 1); ++part) {
          ss[part] += ph[part] * vs;
          tt[part] += ph[part] * vt;
          if (deriv) {
            for (int c = 0; c < 3; ++c) {
              dss[part][c] += ph[part] * vds[c];
              dtt[part][c] += ph[part] * vdt[c];
            }
          }
        }
      }
#pragma unroll
      for (int part = 0; part < (CPLX ? 2 : 1); ++part) {
        s[part] += fac * ss[part];
        t[part] += -0.5 * fac * tt[part];
        if (deriv) {
          for (int c = 0; c < 3; ++c) {
            ds[part][c] += fac * dss[part][c];
            dt[part][c] += -0.5 * fac * dtt[part][c];
          }
        }
      }
    }
  }
  const long nao = (long)n * nloc;
  const long e = ((long)b * nao + (long)A * nloc + o) * nao + (long)Bat * nloc + q;
  // mirror (Bat,q | A,o): both functions are real, so the integral is symmetric and
  // the lattice sum turns into its complex conjugate; the derivative moves to the
  // other centre, which flips its sign (the integrals depend on R_A - R_B only)
  const long em = ((long)b * nao + (long)Bat * nloc + q) * nao + (long)A * nloc + o;
  if (em == e) {
    // diagonal element: written once (its derivative vanishes up to rounding)
    if (!CPLX) {
      S[e] = s[0];
      Tm[e] = t[0];
    } else {
      S[2 * e] = s[0];
      S[2 * e + 1] = s[1];
      Tm[2 * e] = t[0];
      Tm[2 * e + 1] = t[1];
    }
    if (deriv) {
      for (int c = 0; c < 3; ++c) {
        for (int part = 0; part < (CPLX ? 2 : 1); ++part) {
          dS[(CPLX ? 2 : 1) * (e * 3 + c) + part] = ds[part][c];
          dT[(CPLX ? 2 : 1) * (e * 3 + c) + part] = dt[part][c];
        }
      }
    }
    return;
  }
  if (!CPLX) {
    S[e] = s[0];
    Tm[e] = t[0];
    S[em] = s[0];
    Tm[em] = t[0];
    if (deriv) {
      for (int c = 0; c < 3; ++c) {
        dS[e * 3 + c] = ds[0][c];
        dT[e * 3 + c] = dt[0][c];
        dS[em * 3 + c] = -ds[0][c];
        dT[em * 3 + c] = -dt[0][c];
      }
    }
  } else {
    S[2 * e] = s[0];
    S[2 * e + 1] = s[1];
    Tm[2 * e] = t[0];
    Tm[2 * e + 1] = t[1];
    S[2 * em] = s[0];
    S[2 * em + 1] = -s[1];
    Tm[2 * em] = t[0];
    Tm[2 * em + 1] = -t[1];
    if (deriv) {
      for (int c = 0; c < 3; ++c) {
        dS[2 * (e * 3 + c)] = ds[0][c];
        dS[2 * (e * 3 + c) + 1] = ds[1][c];
        dT[2 * (e * 3 + c)] = dt[0][c];
        dT[2 * (e * 3 + c) + 1] = dt[1][c];
        dS[2 * (em * 3 + c)] = -ds[0][c];
        dS[2 * (em * 3 + c) + 1] = ds[1][c];
        dT[2 * (em * 3 + c)] = -dt[0][c];
        dT[2 * (em * 3 + c) + 1] = dt[1][c];
      }
    }
  }
}

// s-only bases (every comp = -1: gth-szv, gth-dzv) at Gamma: one thread per
// unordered atom pair (b, A <= Bat) computes every contracted pair (o, q) of the
// two atoms, so the 4 x 4 x nT lattice exponentials are formed once per atom pair
// instead of once per (o, q); sizes are compile-time, so the Hermite factors stay
// in registers.  Every entry is summed in the order of k_overlap_kinetic (lattice
// terms per primitive pair, then the primitive pairs); the two kernels agree to
// rounding (here the compiler fuses some products into their sums).
template <int NA, int NLOC, bool DERIV>
__global__ __launch_bounds__(256) void k_overlap_kinetic_s(
    const double* __restrict__ xp, long total, int n, const double* __restrict__ lat, int nT,
    const double* __restrict__ alpha, const double* __restrict__ coef, double* __restrict__ S,
    double* __restrict__ Tm, double* __restrict__ dS, double* __restrict__ dT) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= total) return;
  const long npair = (long)n * (n + 1) / 2;
  const int pr = (int)(idx % npair);
  const int b = (int)(idx / npair);
  int A = (int)((2.0 * n + 1.0 - sqrt((2.0 * n + 1.0) * (2.0 * n + 1.0) - 8.0 * pr)) * 0.5);
  if (A < 0) A = 0;
  if (A > n - 1) A = n - 1;
  while (A > 0 && A * n - A * (A - 1) / 2 > pr) --A;
  while (A < n - 1 && (A + 1) * n - (A + 1) * A / 2 <= pr) ++A;
  const int Bat = A + (pr - (A * n - A * (A - 1) / 2));
  const double* xA = xp + ((long)b * n + A) * 3;
  const double* xB = xp + ((long)b * n + Bat) * 3;
  const double d0 = xA[0] - xB[0], d1 = xA[1] - xB[1], d2 = xA[2] - xB[2];
  constexpr int NM = 2 + (DERIV ? 1 : 0);
  double s[NLOC][NLOC], t[NLOC][NLOC], ds[NLOC][NLOC][3], dt[NLOC][NLOC][3];
#pragma unroll
  for (int o = 0; o < NLOC; ++o)
#pragma unroll
    for (int q = 0; q < NLOC; ++q) {
      s[o][q] = 0.0;
      t[o][q] = 0.0;
#pragma unroll
      for (int c = 0; c < 3; ++c) ds[o][q][c] = dt[o][q][c] = 0.0;
    }
  for (int ia = 0; ia < NA; ++ia) {
    for (int ib = 0; ib < NA; ++ib) {
      const double a = alpha[ia], bb = alpha[ib];
      const double mu = a * bb / (a + bb);
      double ss = 0.0, tt = 0.0, dss[3] = {0.0, 0.0, 0.0}, dtt[3] = {0.0, 0.0, 0.0};
      for (int it = 0; it < nT; ++it) {
        const double* T3 = lat + 3 * it;
        const double R[3] = {d0 + T3[0], d1 + T3[1], d2 + T3[2]};
        const double mr2 = mu * (R[0] * R[0] + R[1] * R[1] + R[2] * R[2]);
        if (mr2 > INT_EXP_CUT) continue;
        const double E = exp(-mr2);
        double h[3][NM + 1];
        hermite(mu, R[0], h[0], NM);
        hermite(mu, R[1], h[1], NM);
        hermite(mu, R[2], h[2], NM);
        ss += E;
        tt += E * ((h[0][2] + h[1][2]) + h[2][2]);
        if (DERIV) {
          dss[0] += E * h[0][1];
          dss[1] += E * h[1][1];
          dss[2] += E * h[2][1];
          dtt[0] += E * ((h[0][3] + h[0][1] * h[1][2]) + h[0][1] * h[2][2]);
          dtt[1] += E * ((h[0][2] * h[1][1] + h[1][3]) + h[1][1] * h[2][2]);
          dtt[2] += E * ((h[0][2] * h[2][1] + h[1][2] * h[2][1]) + h[2][3]);
        }
      }
#pragma unroll
      for (int o = 0; o < NLOC; ++o) {
        const double wa = coef[ia * NLOC + o];
        if (wa == 0.0) continue;
#pragma unroll
        for (int q = 0; q < NLOC; ++q) {
          const double wb = coef[ib * NLOC + q];
          if (wb == 0.0) continue;
          const double fac = wa * wb * M_PI * sqrt(M_PI) * pow(a + bb, -1.5);
          s[o][q] += fac * ss;
          t[o][q] += -0.5 * fac * tt;
          if (DERIV) {
#pragma unroll
            for (int c = 0; c < 3; ++c) {
              ds[o][q][c] += fac * dss[c];
              dt[o][q][c] += -0.5 * fac * dtt[c];
            }
          }
        }
      }
    }
  }
  const long nao = (long)n * NLOC;
#pragma unroll
  for (int o = 0; o < NLOC; ++o)
#pragma unroll
    for (int q = 0; q < NLOC; ++q) {
      if (Bat == A && q < o) continue;
      const long e = ((long)b * nao + (long)A * NLOC + o) * nao + (long)Bat * NLOC + q;
      const long em = ((long)b * nao + (long)Bat * NLOC + q) * nao + (long)A * NLOC + o;
      S[e] = s[o][q];
      Tm[e] = t[o][q];
      if (DERIV)
        for (int c = 0; c < 3; ++c) {
          dS[e * 3 + c] = ds[o][q][c];
          dT[e * 3 + c] = dt[o][q][c];
        }
      if (em == e) continue;
      S[em] = s[o][q];
      Tm[em] = t[o][q];
      if (DERIV)
        for (int c = 0; c < 3; ++c) {
          dS[em * 3 + c] = -ds[o][q][c];
          dT[em * 3 + c] = -dt[o][q][c];
        }
    }
}

template <bool CPLX>
static int launch_ovlp(const double* xp, const double* kpt, int B, int n, const double* lattice,
                       int nT, const double* alpha, int na, const double* coef, const int* comp,
                       int nloc, double* S, double* T, double* dS, double* dT, void* stream) {
  if (B < 0 || n <= 0 || nT <= 0 || na <= 0 || nloc <= 0) return HQC_E_ARG;
  if ((dS == nullptr) != (dT == nullptr)) return HQC_E_ARG;
  if (CPLX && !kpt) return HQC_E_ARG;
  if (na > INT_MAX_NA || nloc > INT_MAX_NLOC) return HQC_E_SIZE;
  if (B == 0) return HQC_OK;
  const long total = (long)B * (n * (n + 1) / 2) * nloc * nloc;
  hipLaunchKernelGGL(k_overlap_kinetic<CPLX>, dim3(hqc_div_up(total, 256)), dim3(256), 0,
                     (hipStream_t)stream, xp, kpt, total, n, lattice, nT, alpha, na, coef, comp,
                     nloc, S, T, dS, dT);
  HQC_CHECK_LAUNCH();
  return HQC_OK;
}

extern "C" int hqc_overlap_kinetic(const double* xp, int B, int n, const double* lattice, int nT,
                                   const double* alpha, int na, const double* coef,
                                   const int* comp, int nloc, double* S, double* T, double* dS,
                                   double* dT, void* stream) {
  return launch_ovlp<false>(xp, nullptr, B, n, lattice, nT, alpha, na, coef, comp, nloc, S, T,
                            dS, dT, stream);
}

extern "C" int hqc_overlap_kinetic_k(const double* xp, const double* kpt, int B, int n,
                                     const double* lattice, int nT, const double* alpha, int na,
                                     const double* coef, const int* comp, int nloc, double* S,
                                     double* T, double* dS, double* dT, void* stream) {
  return launch_ovlp<true>(xp, kpt, B, n, lattice, nT, alpha, na, coef, comp, nloc, S, T, dS,
                           dT, stream);
}

extern "C" int hqc_overlap_kinetic_s(const double* xp, int B, int n, const double* lattice, int nT,
                                     const double* alpha, int na, const double* coef, int nloc,
                                     double* S, double* T, double* dS, double* dT, void* stream) {
  if (B < 0 || n <= 0 || nT <= 0 || na <= 0 || nloc <= 0) return HQC_E_ARG;
  if ((dS == nullptr) != (dT == nullptr)) return HQC_E_ARG;
  if (na != 4 || nloc > 2) return HQC_E_SIZE;
  if (B == 0) return HQC_OK;
  const long pairs = (long)B * (n * (n + 1) / 2);
  const dim3 grid(hqc_div_up(pairs, 256));
  const bool der = dS != nullptr;
  hipStream_t st = (hipStream_t)stream;
#define OVLP_S(NL, D)                                                                           \
  hipLaunchKernelGGL((k_overlap_kinetic_s<4, NL, D>), grid, dim3(256), 0, st, xp, pairs, n, lattice, \
                     nT, alpha, coef, S, T, dS, dT)
  if (nloc == 1 && !der) OVLP_S(1, false);
  else if (nloc == 1) OVLP_S(1, true);
  else if (!der) OVLP_S(2, false);
  else OVLP_S(2, true);
#undef OVLP_S
  HQC_CHECK_LAUNCH();
  return HQC_OK;
}
