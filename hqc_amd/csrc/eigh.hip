// Batched symmetric eigensolver for the small FP64 Roothaan problems of the
// SCF (n_ao <= 110): cyclic two-sided Jacobi, one 1024-thread block per
// matrix, everything in LDS.
//
// Replaces jnp.linalg.eigh of the orthogonalised Fock matrix
// (hqc/pbc/scf.py:69,138,198; hqc/pbc/solver.py:741).  Eigenvalues are
// returned ascending with eigenvectors as columns, like eigh.
//
// Round-robin ordering: in every round the n_pad indices (n rounded up to
// even; the extra index is inert) form n_pad/2 disjoint pairs (p, q); all
// pairs are rotated at once.  With R_I = [[c, s], [-s, c]] acting on pair I,
// the new matrix is J^T A J, i.e. every 2x2 block A[{pI,qI},{pJ,qJ}] becomes
// R_I^T A_IJ R_J and every V[i,{pJ,qJ}] becomes V[i,.] R_J; blocks are
// disjoint, so one barrier separates the rotation angles from the update.
// A is kept as its packed upper triangle, V in full, both in LDS.
#include "common.h"

#define EIG_MAXN 110

__device__ __forceinline__ int pk(int i, int j, int n) {
  // packed upper triangle index (i <= j)
  return i * n - (i * (i - 1)) / 2 + (j - i);
}

__device__ __forceinline__ double getA(const double* Ap, int i, int j, int n) {
  return i <= j ? Ap[pk(i, j, n)] : Ap[pk(j, i, n)];
}

#define EIG_THREADS 1024

__global__ __launch_bounds__(EIG_THREADS) void k_jacobi_eigh(const double* __restrict__ Ain, int n,
                                                     int npad, const int* __restrict__ pairs,
                                                     int max_sweeps, double tol,
                                                     double* __restrict__ w,
                                                     double* __restrict__ Vout,
                                                     int* __restrict__ sweeps_out) {
  extern __shared__ double sm[];
  const int np = npad / 2;
  double* Ap = sm;                                // n(n+1)/2
  double* V = Ap + n * (n + 1) / 2;               // n*n
  double* cs = V + n * n;                         // 2*np: (c, s) per pair
  double* red = cs + 2 * np;                      // EIG_THREADS/64 partial sums
  int* pr = reinterpret_cast<int*>(red + EIG_THREADS / 64);  // 2*np: current pairing
  unsigned short* blk = reinterpret_cast<unsigned short*>(pr + 2 * np);  // (I,J) of each block
  const int nblk = np * (np + 1) / 2;
  for (int e = threadIdx.x; e < np; e += blockDim.x) {
    // blocks of row I start at I*np - I*(I-1)/2
    const int start = e * np - (e * (e - 1)) / 2;
    for (int J = e; J < np; ++J) blk[start + (J - e)] = (unsigned short)((e << 8) | J);
  }
  const int b = blockIdx.x;
  const double* A = Ain + (long)b * n * n;
  const int tid = threadIdx.x;
  for (int e = tid; e < n * n; e += blockDim.x) {
    const int i = e / n, j = e - i * n;
    if (i <= j) Ap[pk(i, j, n)] = 0.5 * (A[i * n + j] + A[j * n + i]);
    V[e] = (i == j) ? 1.0 : 0.0;          // column-major: V[col*n + row]
  }
  __syncthreads();
  // scale for the stopping test
  double loc = 0.0;
  for (int e = tid; e < n * (n + 1) / 2; e += blockDim.x) loc += Ap[e] * Ap[e];
  const double fro2 = block_sum(loc, red);
  __shared__ double s_fro2;
  __shared__ int s_done;
  if (tid == 0) s_fro2 = fro2;
  __syncthreads();
  // threshold Jacobi: rotations with |a_pq| <= thr are skipped.  If every
  // off-diagonal element is below thr, sum_{p<q} a_pq^2 < n^2/2 thr^2 =
  // tol^2 |A|_F^2 / 2 and the stopping test passes, so the skip never stalls it
  const double thr = tol * sqrt(s_fro2) / n;
  // static split of the V update: vstep threads per row (blockDim >= n)
  const int vstep = blockDim.x / n;
  const int vrow = vstep > 0 ? tid % n : n;
  const int vj0 = vstep > 0 ? tid / n : 0;
  const bool vrow_ok = vstep > 0 && tid < vstep * n;
  int sweep = 0;
  for (; sweep < max_sweeps; ++sweep) {
    // off-diagonal norm
    double off = 0.0;
    for (int e = tid; e < n * n; e += blockDim.x) {
      const int i = e / n, j = e - i * n;
      if (i < j) {
        const double a = Ap[pk(i, j, n)];
        off += a * a;
      }
    }
    off = block_sum(off, red);
    if (tid == 0) s_done = (off <= tol * tol * s_fro2) ? 1 : 0;
    __syncthreads();
    if (s_done) break;
    for (int r = 0; r < npad - 1; ++r) {
      const int* prow = pairs ? pairs + r * npad : nullptr;
      for (int k = tid; k < np; k += blockDim.x) {
        int p, q;
        if (prow) {
          p = prow[2 * k];
          q = prow[2 * k + 1];
        } else {
          // circle method: index 0 fixed, the others rotated r places; pair k joins
          // positions k and npad-1-k (the schedule hqc_amd's host code builds)
          const int m = npad - 1;
          auto at = [&](int j) { return j == 0 ? 0 : 1 + ((j - 1 - r) % m + m) % m; };
          p = at(k);
          q = at(npad - 1 - k);
        }
        if (p > q) {
          const int t = p;
          p = q;
          q = t;
        }
        pr[2 * k] = p;
        pr[2 * k + 1] = q;
        double c = 1.0, s = 0.0;
        if (q < n) {
          const double apq = Ap[pk(p, q, n)];
          if (fabs(apq) > thr) {
            const double app = Ap[pk(p, p, n)], aqq = Ap[pk(q, q, n)];
            const double theta = (aqq - app) / (2.0 * apq);
            const double t = (theta >= 0.0 ? 1.0 : -1.0) /
                             (fabs(theta) + sqrt(theta * theta + 1.0));
            c = 1.0 / sqrt(t * t + 1.0);
            s = t * c;
          }
        }
        cs[2 * k] = c;
        cs[2 * k + 1] = s;
      }
      __syncthreads();
      // A <- J^T A J by 2x2 blocks (I <= J over pairs)
      for (int e = tid; e < nblk; e += blockDim.x) {
        const int I = blk[e] >> 8, J = blk[e] & 255;
        const int pI = pr[2 * I], qI = pr[2 * I + 1], pJ = pr[2 * J], qJ = pr[2 * J + 1];
        const double cI = cs[2 * I], sI = cs[2 * I + 1], cJ = cs[2 * J], sJ = cs[2 * J + 1];
        const bool vqI = qI < n, vqJ = qJ < n;
        if (I == J) {
          if (!vqI) continue;
          const double app = Ap[pk(pI, pI, n)], aqq = Ap[pk(qI, qI, n)], apq = Ap[pk(pI, qI, n)];
          if (sI != 0.0) {
            const double t = sI / cI;
            Ap[pk(pI, pI, n)] = app - t * apq;
            Ap[pk(qI, qI, n)] = aqq + t * apq;
            Ap[pk(pI, qI, n)] = 0.0;
          }
          continue;
        }
        if (sI == 0.0 && sJ == 0.0) continue;     // both rotations skipped: block unchanged
        // generic block: rows {pI,qI}, cols {pJ,qJ}; M' = R_I^T M R_J
        const double m00 = getA(Ap, pI, pJ, n);
        const double m01 = vqJ ? getA(Ap, pI, qJ, n) : 0.0;
        const double m10 = vqI ? getA(Ap, qI, pJ, n) : 0.0;
        const double m11 = (vqI && vqJ) ? getA(Ap, qI, qJ, n) : 0.0;
        // rows: [x; y] -> [c x - s y; s x + c y]
        const double r00 = cI * m00 - sI * m10, r01 = cI * m01 - sI * m11;
        const double r10 = sI * m00 + cI * m10, r11 = sI * m01 + cI * m11;
        // cols: [u v] -> [c u - s v, s u + c v]
        const double n00 = cJ * r00 - sJ * r01, n01 = sJ * r00 + cJ * r01;
        const double n10 = cJ * r10 - sJ * r11, n11 = sJ * r10 + cJ * r11;
        auto put = [&](int i, int j, double val) {
          if (i <= j) Ap[pk(i, j, n)] = val;
          else Ap[pk(j, i, n)] = val;
        };
        put(pI, pJ, n00);
        if (vqJ) put(pI, qJ, n01);
        if (vqI) put(qI, pJ, n10);
        if (vqI && vqJ) put(qI, qJ, n11);
      }
      // V <- V J: thread owns row vrow and every vstep-th pair from vj0
      if (vrow_ok) {
        for (int J = vj0; J < np; J += vstep) {
          const int pJ = pr[2 * J], qJ = pr[2 * J + 1];
          if (qJ >= n) continue;
          const double c = cs[2 * J], s = cs[2 * J + 1];
          if (s == 0.0) continue;
          const double x = V[pJ * n + vrow], y = V[qJ * n + vrow];
          V[pJ * n + vrow] = c * x - s * y;
          V[qJ * n + vrow] = s * x + c * y;
        }
      }
      __syncthreads();
    }
  }
  // ascending order by rank (ties broken by index)
  for (int i = tid; i < n; i += blockDim.x) {
    const double wi = Ap[pk(i, i, n)];
    int rank = 0;
    for (int j = 0; j < n; ++j) {
      const double wj = Ap[pk(j, j, n)];
      rank += (wj < wi) || (wj == wi && j < i);
    }
    w[(long)b * n + rank] = wi;
    for (int row = 0; row < n; ++row) Vout[((long)b * n + row) * n + rank] = V[i * n + row];
  }
  if (tid == 0 && sweeps_out) sweeps_out[b] = sweep;
}

extern "C" size_t hqc_eigh_lds_bytes(int n) {
  const int npad = n + (n & 1);
  const int np = npad / 2;
  return (size_t)(n * (n + 1) / 2 + n * n + npad + EIG_THREADS / 64) * sizeof(double) +
         (size_t)npad * sizeof(int) + (size_t)(np * (np + 1) / 2) * sizeof(unsigned short);
}

extern "C" int hqc_eigh_jacobi(const double* A, int B, int n, const int* pairs, int max_sweeps,
                               double tol, double* w, double* V, int* sweeps, void* stream) {
  if (B < 0 || n <= 0 || max_sweeps < 0) return HQC_E_ARG;
  if (n > EIG_MAXN) return HQC_E_SIZE;
  if (B == 0) return HQC_OK;
  const int npad = n + (n & 1);
  const size_t lds = hqc_eigh_lds_bytes(n);
  if (lds > 64 * 1024) {
    // raise the dynamic-LDS limit where the runtime asks for it; a runtime that
    // does not support the attribute leaves the launch check below to decide
    if (hipFuncSetAttribute((const void*)k_jacobi_eigh,
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) != hipSuccess)
      (void)hipGetLastError();
  }
  hipLaunchKernelGGL(k_jacobi_eigh, dim3(B), dim3(EIG_THREADS), lds, (hipStream_t)stream, A, n, npad,
                     pairs, max_sweeps, tol, w, V, sweeps);
  HQC_CHECK_LAUNCH();
  return HQC_OK;
}
