// Orbital-form density rho(r) = sum_i (sum_mu At[mu][i] phi_mu(r))^2 from int8
// slices with exact integer accumulation on the int8 matrix cores
// (v_mfma_i32_16x16x64_i8).  It serves the mixed path's accurate density; the
// reference evaluates the same quantity in the D form, rho = einsum('pr,qr,pq->r',
// phi, phi, D) with D = C f C^T (hqc/pbc/solver.py:690); here At = C sqrt(f).
//
// Slicing.  A value x of a vector whose largest magnitude is below 2^e is cut
// into int8 digits of base 256 by
//     y_0 = x 2^(6-e),  x_s = rint(y_s),  y_(s+1) = 256 (y_s - x_s),
// every step exact in FP64, |x_0| <= 64 and |x_s| <= 128; a digit +128 (only
// when y_s - x_s was exactly 1/2, so x_(s-1) is even and below 127) becomes -128
// with a carry of 1 into x_(s-1), which keeps every digit in [-128, 127] and the
// value unchanged:
//     x = 2^(e-6) sum_s 2^(-8 s) x_s + r,   |r| <= 2^(e-7-8(S-1))  for S digits.
// phi is sliced once per geometry with a per-grid-point exponent (S_phi digits,
// a fixed perturbation of the basis: smooth in the SCF state); the orbital
// coefficients are sliced per call with a per-orbital exponent into T+1 digits,
// and every digit product with k + l <= T is accumulated exactly in int32 per
// level d = k + l (|level| <= 4 * 128 * 128 * 128 < 2^31).  What changes with the
// SCF state from call to call (coefficient rounding, dropped levels) is then
// ~2^(-6-8T) of max|At_i| max|phi(r)| per AO: 2^-46 with T = 5, 2^-38 with T = 4
// (the SCF's early iterations).  psi_i(r) = 2^(e_r + e_i - 12) sum_d 2^(-8 d) P_d,
// combined in FP64 by Horner's rule.
//
// Layout of the phi slices (per geometry, ld/16 tiles of 16 grid points):
//     q[tile][s][mb][16 points][16 AOs]   (NB = nao_pad/16 AO blocks mb),
// so a lane's 16-byte operand (16 consecutive AOs of one point) is one load and
// a wave's 64 lanes read 1 KB contiguously; pe[point] holds the exponent e_r.
#include "common.h"

typedef int i32x4 __attribute__((ext_vector_type(4)));
typedef unsigned u32x4q __attribute__((ext_vector_type(4)));

#define Q8_TMAX 5         // highest digit level kept at most: T + 1 coefficient digits
#define Q8_MAXKS 2        // nao_pad <= 128 (two K steps of 64 AOs)
#define Q8_NW 4           // waves per workgroup
// point tiles per wave at least: the coefficient digits are cut once per
// workgroup, so a few geometries spread over too many workgroups pay that
// prologue over and over (4 geometries: 0.117 -> 0.096 ms)
#define Q8_MIN_TILES 8
#define Q8_TARGET_BLOCKS 2048

__device__ __forceinline__ int q8_exponent(double m) {
  // smallest e with m < 2^e (0 for m = 0)
  return m > 0.0 ? ilogb(m) + 1 : 0;
}

// the N base-256 digits of y = x 2^(6-e) (|y| < 64), each in [-128, 127]
template <int N>
__device__ __forceinline__ void q8_digits(double y, int (&dg)[N]) {
#pragma unroll
  for (int s = 0; s < N; ++s) {
    const double d = rint(y);
    dg[s] = (int)d;
    y = (y - d) * 256.0;
  }
#pragma unroll
  for (int s = N - 1; s > 0; --s) {
    const int c = dg[s] == 128;
    dg[s] -= 256 * c;
    dg[s - 1] += c;
  }
}

// ---------------------------------------------------------------------------
// phi [G][nao_pad][ld] (FP64) -> q (layout above) and pe [G][ld]
// grid (ld/64, G), 256 threads: point p = 64 blockIdx.x + lane, AO blocks
// mb = w, w + 4 of that point per thread.  amax [G] (optional, zeroed by the
// caller) receives max |phi| of each geometry: the workgroup's largest point
// maximum enters by one 64-bit atomic max on its bit pattern (non-negative
// doubles order as their bits), which saves the caller a separate pass over phi.
// H2: the same pass also writes the FP16 (hi, lo) pairs of scale*phi in phi's
// layout, the values hqc_phi_split_f16 gives (one read of phi for both forms).
// ---------------------------------------------------------------------------
__device__ __forceinline__ unsigned f16_pair_bits(double x) {
  const _Float16 hi = (_Float16)x;
  const _Float16 lo = (_Float16)(x - (double)hi);
  return (unsigned)__builtin_bit_cast(unsigned short, hi) |
         ((unsigned)__builtin_bit_cast(unsigned short, lo) << 16);
}

template <int SP, bool H2>
__global__ __launch_bounds__(256) void k_phi_slice_i8(const double* __restrict__ phi, int nao_pad,
                                                      int ld, signed char* __restrict__ q,
                                                      short* __restrict__ pe,
                                                      unsigned long long* __restrict__ amax,
                                                      double scale, unsigned* __restrict__ h2) {
  __shared__ double red[4][64];
  const int NB = nao_pad >> 4;
  const int g = blockIdx.y;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const long p = (long)blockIdx.x * 64 + lane;
  const double* P = phi + (long)g * nao_pad * ld + p;
  double v[2][16];
  double m = 0.0;
#pragma unroll
  for (int t = 0; t < 2; ++t) {
    const int mb = w + 4 * t;
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      v[t][i] = mb < NB ? P[(long)(16 * mb + i) * ld] : 0.0;
      m = fmax(m, fabs(v[t][i]));
    }
  }
  if (H2) {
    unsigned* Hp = h2 + (long)g * nao_pad * ld + p;
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      const int mb = w + 4 * t;
      if (mb >= NB) continue;
#pragma unroll
      for (int i = 0; i < 16; ++i) Hp[(long)(16 * mb + i) * ld] = f16_pair_bits(v[t][i] * scale);
    }
  }
  red[w][lane] = m;
  __syncthreads();
  m = fmax(fmax(red[0][lane], red[1][lane]), fmax(red[2][lane], red[3][lane]));
  const int e = q8_exponent(m);
  if (w == 0) {
    pe[(long)g * ld + p] = (short)e;
    if (amax) {
      double mw = m;
      for (int o = 32; o > 0; o >>= 1) mw = fmax(mw, __shfl_xor(mw, o, 64));
      if (lane == 0) atomicMax(amax + g, (unsigned long long)__double_as_longlong(mw));
    }
  }
  signed char* Q = q + ((long)g * ld + (p & ~15L)) * SP * nao_pad + (p & 15) * 16;
#pragma unroll
  for (int t = 0; t < 2; ++t) {
    const int mb = w + 4 * t;
    if (mb >= NB) continue;
    unsigned word[SP][4];
#pragma unroll
    for (int s = 0; s < SP; ++s)
#pragma unroll
      for (int c = 0; c < 4; ++c) word[s][c] = 0u;
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      int dg[SP];
      q8_digits(ldexp(v[t][i], 6 - e), dg);
#pragma unroll
      for (int s = 0; s < SP; ++s) word[s][i >> 2] |= ((unsigned)dg[s] & 0xffu) << (8 * (i & 3));
    }
#pragma unroll
    for (int s = 0; s < SP; ++s)
      *reinterpret_cast<u32x4q*>(Q + ((long)s * NB + mb) * 256) =
          u32x4q{word[s][0], word[s][1], word[s][2], word[s][3]};
  }
}

// ---------------------------------------------------------------------------
// rho[b][r] = sum_i psi_i(r)^2 over the m_pad columns of At[b] ([nao_pad][m_pad]).
// grid (nwg, B), Q8_NW waves; per group of up to 64 orbitals the workgroup
// slices At into LDS (digit-major, rows of 64 KS + 16 bytes so the 16 lanes of
// a fragment read hit distinct banks), then every wave walks its own tiles of
// 16 grid points two at a time, each pair through MT orbital tiles x KS K-steps
// x the kept digit pairs with every coefficient fragment read once for both.
// ---------------------------------------------------------------------------
// KS = 2: two waves per SIMD, which is what the LDS allows (two workgroups per
// CU); left to itself the compiler gave SP = 3 260 registers and one wave
// (3.89 -> 3.33 ms for 256 geometries, 62 orbitals, 112 AOs)
template <int SP, int KS, int T>
__global__ __launch_bounds__(64 * Q8_NW) __attribute__((amdgpu_waves_per_eu(KS))) void
k_orb_density_q8(
    const signed char* __restrict__ q, const short* __restrict__ pe, const int* __restrict__ bidx,
    const double* __restrict__ At, int nao_pad, int m_pad, int ld, int spt,
    double* __restrict__ rho) {
  constexpr int STRA = 64 * KS + 16;
  constexpr int PLANE = 64 * STRA;              // one digit of 64 orbitals
  constexpr int SC = T + 1;                     // coefficient digits
  extern __shared__ unsigned char smem_q8[];
  signed char* As = reinterpret_cast<signed char*>(smem_q8);          // [SC][64][STRA]
  double* wo = reinterpret_cast<double*>(smem_q8 + SC * PLANE);      // [64] 2^(2 e_i)
  double* red = wo + 64;                                              // [4][64]
  const int NB = nao_pad >> 4;
  const int NPT = ld >> 4;
  const int b = blockIdx.y;
  const long gb = bidx ? bidx[b] : b;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int j = lane & 15, kq = lane >> 4;
  const double* Ab = At + (long)b * nao_pad * m_pad;
  // spt digits are stored per value; the first SP of them are used
  const signed char* Qb = q + gb * (long)ld * spt * nao_pad + j * 16;
  const short* Pe = pe + gb * ld;
  double* Rb = rho + (long)b * ld;
  const int wstride = gridDim.x * Q8_NW;
  const int pt0 = blockIdx.x * Q8_NW + w;

  // AO blocks past nao_pad re-read the last block: their coefficient digits are
  // zero, so the (exact integer) products vanish whatever these bytes hold, and
  // the loaded values are not touched before the MFMAs that use them
  auto loadB = [&](int pt, u32x4q (&Bf)[SP][KS], int& er) {
    const signed char* base = Qb + (long)pt * spt * NB * 256;
#pragma unroll
    for (int s = 0; s < SP; ++s)
#pragma unroll
      for (int ks = 0; ks < KS; ++ks)
        Bf[s][ks] = *reinterpret_cast<const u32x4q*>(
            base + ((long)s * NB + min(4 * ks + kq, NB - 1)) * 256);
    er = Pe[pt * 16 + j];
  };

  for (int og = 0; og < m_pad; og += 64) {
    const int MW = min(64, m_pad - og);         // a multiple of 16
    __syncthreads();                            // previous group's fragments are read
    {
      // per-orbital exponent: orbital i = lane, AO rows w, w + 4, ...
      double m = 0.0;
      if (lane < MW)
        for (int mu = w; mu < nao_pad; mu += 4) m = fmax(m, fabs(Ab[(long)mu * m_pad + og + lane]));
      red[w * 64 + lane] = m;
    }
    __syncthreads();
    {
      const double m = fmax(fmax(red[lane], red[64 + lane]), fmax(red[128 + lane], red[192 + lane]));
      const int e = q8_exponent(m);
      if (w == 0) wo[lane] = ldexp(1.0, 2 * e);
      // digits of 4 consecutive AOs per task: orbital i = lane, AO quad from w
      for (int quad = w; quad < 16 * KS; quad += 4) {
        unsigned word[SC];
#pragma unroll
        for (int s = 0; s < SC; ++s) word[s] = 0u;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          const int mu = 4 * quad + c;
          int dg[SC];
          q8_digits((lane < MW && mu < nao_pad) ? ldexp(Ab[(long)mu * m_pad + og + lane], 6 - e)
                                                 : 0.0, dg);
#pragma unroll
          for (int s = 0; s < SC; ++s) word[s] |= ((unsigned)dg[s] & 0xffu) << (8 * c);
        }
#pragma unroll
        for (int s = 0; s < SC; ++s)
          *reinterpret_cast<unsigned*>(As + s * PLANE + lane * STRA + 4 * quad) = word[s];
      }
    }
    __syncthreads();
    const int MT = MW >> 4;
    // point tiles in pairs: each coefficient fragment read from LDS serves both
    // (one pair buffer; 3.74 ms against 3.98 for single tiles with the next
    // tile's loads in flight, while two pair buffers cost occupancy)
    auto tile2 = [&](const u32x4q (&Ba)[SP][KS], const u32x4q (&Bb)[SP][KS], int era, int erb,
                     int pta, int ptb) {
      double ra = 0.0, rb = 0.0;
      for (int t = 0; t < MT; ++t) {
        i32x4 acca[T + 1], accb[T + 1];
#pragma unroll
        for (int d = 0; d <= T; ++d) {
          acca[d] = i32x4{0, 0, 0, 0};
          accb[d] = i32x4{0, 0, 0, 0};
        }
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) {
          i32x4 a[SC];
#pragma unroll
          for (int k = 0; k < SC; ++k)
            a[k] = *reinterpret_cast<const i32x4*>(As + k * PLANE + (16 * t + j) * STRA + 64 * ks +
                                                   16 * kq);
#pragma unroll
          for (int l = 0; l < SP; ++l) {
            const i32x4 bla = __builtin_bit_cast(i32x4, Ba[l][ks]);
            const i32x4 blb = __builtin_bit_cast(i32x4, Bb[l][ks]);
#pragma unroll
            for (int k = 0; k + l <= T; ++k) {
              acca[k + l] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a[k], bla, acca[k + l], 0, 0, 0);
              accb[k + l] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a[k], blb, accb[k + l], 0, 0, 0);
            }
          }
        }
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          double Qa = (double)acca[T][e], Qb = (double)accb[T][e];
#pragma unroll
          for (int d = T - 1; d >= 0; --d) {
            Qa = fma(Qa, 0x1p-8, (double)acca[d][e]);
            Qb = fma(Qb, 0x1p-8, (double)accb[d][e]);
          }
          const double wv = wo[16 * t + 4 * kq + e];
          ra = fma(wv * Qa, Qa, ra);
          rb = fma(wv * Qb, Qb, rb);
        }
      }
      ra += __shfl_xor(ra, 16, 64);
      ra += __shfl_xor(ra, 32, 64);
      rb += __shfl_xor(rb, 16, 64);
      rb += __shfl_xor(rb, 32, 64);
      if (kq == 0) {
        const double r = ldexp(ra, 2 * era - 24);
        const long idx = (long)pta * 16 + j;
        Rb[idx] = og == 0 ? r : Rb[idx] + r;
        if (ptb < NPT) {                        // the pair's second tile exists
          const double r2 = ldexp(rb, 2 * erb - 24);
          const long idx2 = (long)ptb * 16 + j;
          Rb[idx2] = og == 0 ? r2 : Rb[idx2] + r2;
        }
      }
    };
    u32x4q Ba[SP][KS], Bb[SP][KS];
#pragma unroll
    for (int s2 = 0; s2 < SP; ++s2)
#pragma unroll
      for (int ks = 0; ks < KS; ++ks) Bb[s2][ks] = u32x4q{0u, 0u, 0u, 0u};
    int ea = 0, eb = 0;
    for (int pt = pt0; pt < NPT; pt += 2 * wstride) {   // wave-uniform
      loadB(pt, Ba, ea);
      if (pt + wstride < NPT) loadB(pt + wstride, Bb, eb);
      tile2(Ba, Bb, ea, eb, pt, pt + wstride);
    }
  }
}

static size_t q8_lds_bytes(int KS, int T) {
  return (size_t)(T + 1) * 64 * (64 * KS + 16) + 64 * sizeof(double) + 256 * sizeof(double);
}

static int launch_phi_slice(const double* phi, int G, int nao_pad, int ld, int nslice,
                            signed char* q, short* pe, double* amax, double scale, unsigned* h2,
                            void* stream) {
  if (G < 0 || nao_pad <= 0 || (nao_pad & 15) || ld <= 0 || (ld & 63) || !phi || !q || !pe)
    return HQC_E_ARG;
  if (h2 && !(scale > 0.0)) return HQC_E_ARG;
  if (nao_pad > 64 * Q8_MAXKS) return HQC_E_SIZE;
  if (G == 0) return HQC_OK;
  const dim3 grid(ld / 64, G);
  hipStream_t s = (hipStream_t)stream;
  unsigned long long* am = reinterpret_cast<unsigned long long*>(amax);
#define SLICE_CASE(SP)                                                                        \
  case SP:                                                                                    \
    if (h2)                                                                                   \
      hipLaunchKernelGGL((k_phi_slice_i8<SP, true>), grid, dim3(256), 0, s, phi, nao_pad, ld, \
                         q, pe, am, scale, h2);                                               \
    else                                                                                      \
      hipLaunchKernelGGL((k_phi_slice_i8<SP, false>), grid, dim3(256), 0, s, phi, nao_pad,    \
                         ld, q, pe, am, scale, h2);                                           \
    break;
  switch (nslice) {
    SLICE_CASE(3)
    SLICE_CASE(4)
    SLICE_CASE(5)
    default: return HQC_E_ARG;
  }
#undef SLICE_CASE
  HQC_CHECK_LAUNCH();
  return HQC_OK;
}

extern "C" int hqc_phi_slice_i8(const double* phi, int G, int nao_pad, int ld, int nslice,
                                signed char* q, short* pe, double* amax, void* stream) {
  return launch_phi_slice(phi, G, nao_pad, ld, nslice, q, pe, amax, 1.0, nullptr, stream);
}

extern "C" int hqc_phi_slice_i8_h2(const double* phi, int G, int nao_pad, int ld, int nslice,
                                   double scale, signed char* q, short* pe, unsigned* h2,
                                   void* stream) {
  if (!h2) return HQC_E_ARG;
  return launch_phi_slice(phi, G, nao_pad, ld, nslice, q, pe, nullptr, scale, h2, stream);
}

extern "C" int hqc_orb_density_q8(const signed char* q, const short* pe, const int* bidx,
                                  const double* At, int B, int nao_pad, int m_pad, int ld,
                                  int nslice, int nuse, int top, double* rho, void* stream) {
  if (B < 0 || nao_pad <= 0 || (nao_pad & 15) || m_pad <= 0 || (m_pad & 15) || ld <= 0 ||
      (ld & 63) || nslice < 3 || nslice > 5 || nuse < 3 || nuse > nslice || top < 3 ||
      top > Q8_TMAX || (top == 3 && nuse > 4))
    return HQC_E_ARG;
  if (nao_pad > 64 * Q8_MAXKS) return HQC_E_SIZE;
  if (B == 0) return HQC_OK;
  if (!q || !pe || !At || !rho) return HQC_E_ARG;
  const int KS = nao_pad > 64 ? 2 : 1;
  const size_t lds = q8_lds_bytes(KS, top);
  const int npt = ld >> 4;
  long nwg = hqc_tuned(HQC_TUNE_Q8_BLOCKS, Q8_TARGET_BLOCKS) / B;
  const long min_tiles = hqc_tuned(HQC_TUNE_Q8_MIN_TILES, Q8_MIN_TILES);
  const long maxwg = (npt + Q8_NW * min_tiles - 1) / (Q8_NW * min_tiles);
  if (nwg > maxwg) nwg = maxwg;
  if (nwg < 1) nwg = 1;
  const dim3 grid((unsigned)nwg, B);
  hipStream_t s = (hipStream_t)stream;
#define Q8_CASE(SP_, KS_, T_)                                                                  \
  if (nuse == SP_ && KS == KS_ && top == T_) {                                                 \
    if (lds > 64 * 1024 &&                                                                     \
        hipFuncSetAttribute((const void*)k_orb_density_q8<SP_, KS_, T_>,                       \
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) != hipSuccess) \
      (void)hipGetLastError();                                                                 \
    hipLaunchKernelGGL((k_orb_density_q8<SP_, KS_, T_>), grid, dim3(64 * Q8_NW), lds, s, q, pe, \
                       bidx, At, nao_pad, m_pad, ld, nslice, rho);                             \
  }
  Q8_CASE(3, 1, 3) Q8_CASE(3, 2, 3) Q8_CASE(4, 1, 3) Q8_CASE(4, 2, 3)
  Q8_CASE(3, 1, 4) Q8_CASE(3, 2, 4) Q8_CASE(4, 1, 4) Q8_CASE(4, 2, 4) Q8_CASE(5, 1, 4)
  Q8_CASE(5, 2, 4) Q8_CASE(3, 1, 5) Q8_CASE(3, 2, 5) Q8_CASE(4, 1, 5) Q8_CASE(4, 2, 5)
  Q8_CASE(5, 1, 5) Q8_CASE(5, 2, 5)
#undef Q8_CASE
  HQC_CHECK_LAUNCH();
  return HQC_OK;
}
