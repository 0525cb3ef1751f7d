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

// slot j of the circle-method schedule in round r: index 0 stays, the other
// m = npad - 1 indices move one place per round (the host table's order)
__device__ __forceinline__ int circle_at(int j, int r, int m) {
  if (j == 0) return 0;
  const int x = j - 1 - r;                  // in (-m, m)
  return 1 + (x < 0 ? x + m : x);
}

template <bool TABLE>
__device__ __forceinline__ void round_pair(const int* __restrict__ prow, int k, int r, int npad,
                                           int& p, int& q) {
  int a, b;
  if (TABLE) {
    a = prow[2 * k];
    b = prow[2 * k + 1];
  } else {
    a = circle_at(k, r, npad - 1);
    b = circle_at(npad - 1 - k, r, npad - 1);
  }
  p = min(a, b);
  q = max(a, b);
}

// (c, s) of the rotation that zeroes a_pq: t = tan(angle) is the root of
// t^2 + 2 theta t - 1 = 0 of smaller magnitude, theta = (a_qq - a_pp) / (2 a_pq).
// Only c and s enter the update, and c = 1/sqrt(1 + t^2), s = t c is an exact
// rotation for any t, so t comes from the hardware reciprocal; the rotated block
// is then formed from (c, s) as R^T M R, so A and V stay consistent.
__device__ __forceinline__ double2 jacobi_cs(double app, double aqq, double apq) {
  const double d = aqq - app;
  const double h = sqrt(d * d + 4.0 * apq * apq);
  double t = 2.0 * fabs(apq) * __builtin_amdgcn_rcp(fabs(d) + h);
  if (d != 0.0 && ((d < 0.0) != (apq < 0.0))) t = -t;   // sign of theta
  const double c = rsqrt(1.0 + t * t);
  return make_double2(c, t * c);
}

// FAST: the round's pairs are written to LDS once (with the packed-row offsets of
// both indices) by the threads that form the rotations, so the block and V
// updates read them instead of recomputing positions and triangle offsets; the
// arithmetic on A and V is the same, so are the results.
template <bool TABLE, bool FAST>
__global__ __launch_bounds__(EIG_THREADS) void k_jacobi_eigh(const double* __restrict__ Ain, int n,
                                                     int npad, const int* __restrict__ pairs,
                                                     int max_sweeps, double tol,
                                                     double* __restrict__ w,
                                                     double* __restrict__ Vout,
                                                     int* __restrict__ sweeps_out) {
  extern __shared__ double sm[];
  const int np = npad / 2;
  const int m = npad - 1;                         // rounds per sweep
  const int na = n * (n + 1) / 2;
  double* Ap = sm;                                // n(n+1)/2
  double* V = Ap + na;                            // n*n
  const int cso = (na + n * n + 1) & ~1;          // 16-byte aligned
  double2* cs = reinterpret_cast<double2*>(sm + cso);   // np: (c, s) per pair
  double* red = sm + cso + 2 * np;                // EIG_THREADS/64 partial sums
  int4* prs = reinterpret_cast<int4*>(red + EIG_THREADS / 64);   // np: (p, q, off_p, off_q)
  __shared__ double s_fro2;
  __shared__ int s_done, s_rot;
  const int nblk = np * (np + 1) / 2;
  const int tid = threadIdx.x;
  // this thread's blocks (I <= J) of the pair-by-pair partition, fixed for the
  // whole solve: e = tid and tid + EIG_THREADS in row-major order of the triangle
  int bI[2], bJ[2];
#pragma unroll
  for (int u = 0; u < 2; ++u) {
    const int e = tid + u * EIG_THREADS;
    int I = 0, start = 0;
    while (I < np && start + (np - I) <= e) {     // row I holds blocks (I, I..np-1)
      start += np - I;
      ++I;
    }
    bI[u] = e < nblk ? I : -1;
    bJ[u] = I + (e - start);
  }
  const int b = blockIdx.x;
  const double* A = Ain + (long)b * n * n;
  for (int e = tid; e < n * n; e += blockDim.x) {
    const int i = e / n, j = e - i * n;
    if (i <= j) Ap[pk(i, j, n)] = 0.5 * (A[i * n + j] + A[j * n + i]);
    V[e] = (i == j) ? 1.0 : 0.0;          // column-major: V[col*n + row]
  }
  if (tid == 0) s_rot = 0;
  __syncthreads();
  // scale for the stopping test
  double loc = 0.0;
  for (int e = tid; e < na; e += blockDim.x) loc += Ap[e] * Ap[e];
  const double fro2 = block_sum(loc, red);
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
  const int vG = max(1, (int)blockDim.x / np);      // threads per pair (FAST)
  const int vJ = tid / vG, vr0 = tid % vG;
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
    for (int r = 0; r < m; ++r) {
      const int* prow = TABLE ? pairs + r * npad : nullptr;
      const int gr = sweep * m + r + 1;           // this round's id
      if (tid < np) {
        int p, q;
        round_pair<TABLE>(prow, tid, r, npad, p, q);
        // off_i + j is the packed index of (i, j), i <= j
        if (FAST) prs[tid] = make_int4(p, q, pk(p, 0, n), q < n ? pk(q, 0, n) : 0);
        double2 rot = make_double2(1.0, 0.0);
        if (q < n) {
          const double apq = Ap[pk(p, q, n)];
          if (fabs(apq) > thr) {
            rot = jacobi_cs(Ap[pk(p, p, n)], Ap[pk(q, q, n)], apq);
            s_rot = gr;
          }
        }
        cs[tid] = rot;
      }
      __syncthreads();
      // a round without a rotation leaves A and V as they are (uniform branch:
      // s_rot only changes again after every thread has passed this point)
      if (s_rot != gr) continue;
      // A <- J^T A J by 2x2 blocks (I <= J over pairs)
#pragma unroll
      for (int u = 0; u < 2; ++u) {
        const int I = bI[u], J = bJ[u];
        if (I < 0) continue;
        const double2 rI = cs[I], rJ = cs[J];
        if (rI.y == 0.0 && rJ.y == 0.0) continue;   // both rotations skipped
        int pI, qI, pJ, qJ;
        int oPI = 0, oQI = 0, oPJ = 0, oQJ = 0;
        if (FAST) {
          const int4 a = prs[I], c = prs[J];
          pI = a.x; qI = a.y; oPI = a.z; oQI = a.w;
          pJ = c.x; qJ = c.y; oPJ = c.z; oQJ = c.w;
        } else {
          round_pair<TABLE>(prow, I, r, npad, pI, qI);
          round_pair<TABLE>(prow, J, r, npad, pJ, qJ);
        }
        const bool vqI = qI < n, vqJ = qJ < n;
        const double cI = rI.x, sI = rI.y, cJ = rJ.x, sJ = rJ.y;
        // packed index of (i, j) from the row offsets of both
        auto ix = [&](int i, int oi, int j, int oj) { return i <= j ? oi + j : oj + i; };
        // rows {pI,qI}, cols {pJ,qJ}; M' = R_I^T M R_J (for I == J the
        // symmetric diagonal block, whose off-diagonal comes out ~0)
        double m00, m01, m10, m11;
        if (FAST) {
          m00 = Ap[ix(pI, oPI, pJ, oPJ)];
          m01 = vqJ ? Ap[ix(pI, oPI, qJ, oQJ)] : 0.0;
          m10 = vqI ? Ap[ix(qI, oQI, pJ, oPJ)] : 0.0;
          m11 = (vqI && vqJ) ? Ap[ix(qI, oQI, qJ, oQJ)] : 0.0;
        } else {
          m00 = getA(Ap, pI, pJ, n);
          m01 = vqJ ? getA(Ap, pI, qJ, n) : 0.0;
          m10 = vqI ? getA(Ap, qI, pJ, n) : 0.0;
          m11 = (vqI && vqJ) ? getA(Ap, qI, qJ, n) : 0.0;
        }
        // rows: [x; y] -> [c x - s y; s x + c y]
        const double r00 = cI * m00 - sI * m10, r01 = cI * m01 - sI * m11;
        const double r10 = sI * m00 + cI * m10, r11 = sI * m01 + cI * m11;
        // cols: [u v] -> [c u - s v, s u + c v]
        const double n00 = cJ * r00 - sJ * r01, n01 = sJ * r00 + cJ * r01;
        const double n10 = cJ * r10 - sJ * r11, n11 = sJ * r10 + cJ * r11;
        if (FAST) {
          if (I == J) {
            if (!vqI) continue;
            Ap[oPI + pI] = n00;
            Ap[oQI + qI] = n11;
            Ap[ix(pI, oPI, qI, oQI)] = 0.5 * (n01 + n10);
            continue;
          }
          Ap[ix(pI, oPI, pJ, oPJ)] = n00;
          if (vqJ) Ap[ix(pI, oPI, qJ, oQJ)] = n01;
          if (vqI) Ap[ix(qI, oQI, pJ, oPJ)] = n10;
          if (vqI && vqJ) Ap[ix(qI, oQI, qJ, oQJ)] = n11;
          continue;
        }
        auto put = [&](int i, int j, double val) {
          if (i <= j) Ap[pk(i, j, n)] = val;
          else Ap[pk(j, i, n)] = val;
        };
        if (I == J) {
          if (!vqI) continue;
          Ap[pk(pI, pI, n)] = n00;
          Ap[pk(qI, qI, n)] = n11;
          Ap[pk(pI, qI, n)] = 0.5 * (n01 + n10);
          continue;
        }
        put(pI, pJ, n00);
        if (vqJ) put(pI, qJ, n01);
        if (vqI) put(qI, pJ, n10);
        if (vqI && vqJ) put(qI, qJ, n11);
      }
      // V <- V J.  FAST: thread owns pair vJ and rows vr0, vr0 + vG, ... (one read
      // of its rotation and positions); else row vrow and every vstep-th pair from vj0
      if (FAST) {
        if (vJ < np) {
          const double2 rJ = cs[vJ];
          const int4 c = prs[vJ];
          if (rJ.y != 0.0 && c.y < n) {
            for (int row = vr0; row < n; row += vG) {
              const double x = V[c.x * n + row], y = V[c.y * n + row];
              V[c.x * n + row] = rJ.x * x - rJ.y * y;
              V[c.y * n + row] = rJ.y * x + rJ.x * y;
            }
          }
        }
      } else if (vrow_ok) {
        for (int J = vj0; J < np; J += vstep) {
          const double2 rJ = cs[J];
          if (rJ.y == 0.0) continue;
          int pJ, qJ;
          if (FAST) {
            const int4 c = prs[J];
            pJ = c.x;
            qJ = c.y;
          } else {
            round_pair<TABLE>(prow, J, r, npad, pJ, qJ);
          }
          if (qJ >= n) continue;
          const double x = V[pJ * n + vrow], y = V[qJ * n + vrow];
          V[pJ * n + vrow] = rJ.x * x - rJ.y * y;
          V[qJ * n + vrow] = rJ.y * x + rJ.x * y;
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
  const int cso = (n * (n + 1) / 2 + n * n + 1) & ~1;
  return (size_t)(cso + 2 * np + EIG_THREADS / 64) * sizeof(double) + (size_t)np * sizeof(int4);
}

extern "C" int hqc_eigh_jacobi(const double* A, int B, int n, const int* pairs, int max_sweeps,
                               double tol, double* w, double* V, int* sweeps, void* stream) {
  if (B < 0 || n <= 0 || max_sweeps < 0) return HQC_E_ARG;
  if (n > EIG_MAXN) return HQC_E_SIZE;
  if (B == 0) return HQC_OK;
  const int npad = n + (n & 1);
  const size_t lds = hqc_eigh_lds_bytes(n);
  const bool fast = hqc_tuned(HQC_TUNE_EIGH_RECOMPUTE, 0) == 0;
  const void* kern = pairs ? (fast ? (const void*)k_jacobi_eigh<true, true>
                                   : (const void*)k_jacobi_eigh<true, false>)
                           : (fast ? (const void*)k_jacobi_eigh<false, true>
                                   : (const void*)k_jacobi_eigh<false, false>);
  if (lds > 64 * 1024) {
    // raise the dynamic-LDS limit where the runtime asks for it; a runtime that
    // does not support the attribute leaves the launch check below to decide
    if (hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) !=
        hipSuccess)
      (void)hipGetLastError();
  }
  hipStream_t st = (hipStream_t)stream;
#define EIG_LAUNCH(T_, F_)                                                                     \
  hipLaunchKernelGGL((k_jacobi_eigh<T_, F_>), dim3(B), dim3(EIG_THREADS), lds, st, A, n, npad, \
                     pairs, max_sweeps, tol, w, V, sweeps)
  if (pairs && fast)
    EIG_LAUNCH(true, true);
  else if (pairs)
    EIG_LAUNCH(true, false);
  else if (fast)
    EIG_LAUNCH(false, true);
  else
    EIG_LAUNCH(false, false);
#undef EIG_LAUNCH
  HQC_CHECK_LAUNCH();
  return HQC_OK;
}
