// Batched 3-D complex FFT for the odd (and even) grids of the GPW solver.
//
// When a whole (y, z) plane fits in LDS (g <= 69 with two buffers, g <= 98 in
// place, FP64) the y and z axes are done together on planes (k_fft_plane), then
// the x axis by lines; larger grids go axis by axis.  A line block stages `lpb`
// lines of length g in LDS (position-major, line-minor so that strided axes load
// coalesced) and runs mixed-radix Stockham passes, either between two LDS
// buffers or in place (a thread reads the inputs of its work items into
// registers, the block synchronises, and the outputs are written over them).
//
// Arithmetic.  Every output is formed by the same sequence of explicit FMAs,
// adds and multiplies whichever schedule (buffers, threads, items per thread)
// runs it, so all schedules give the same bits:
//  * radix 2 and 4: adds only (the constants are +-1 and +-i);
//  * odd radix R <= 13: the twiddled inputs v_r are paired,
//        a_r = v_r + v_(R-r),  b_r = v_r - v_(R-r),  r = 1 .. (R-1)/2,
//        X_0 = v_0 + sum_r a_r,
//        X_k = T1 + i s T2,  X_(R-k) = T1 - i s T2,   k = 1 .. (R-1)/2,
//        T1 = v_0 + sum_r a_r cos(2 pi r k / R),  T2 = sum_r b_r sin(2 pi r k / R),
//    s = the transform's sign, with the cosines and sines as compile-time
//    constants: a quarter of the multiplies of the direct R-point sum and no
//    table reads;
//  * any other factor (17, 19, 23, ..., and a prime length > 13 as the only
//    stage): the same pairing in two sub-passes, first (a_r, b_r) over the
//    twiddled inputs in place, then FFT_PAIRS output pairs per work item with
//    cos and s*sin read from the twiddle table.
// Inter-stage twiddles come from a length-g table exp(s 2 pi i m/g) built in
// LDS with sincospi.
#include "common.h"

#define FFT_MAX_G 128
#define FFT_MAX_STAGES 12
#define FFT_THREADS 256
#define FFT_PAIRS 4  // output pairs (X_k, X_(R-k)) per work item, generic factor

struct FftPlan {
  int nst;
  int rad[FFT_MAX_STAGES];
};

// Where grid b of a batch starts: rows of `rows` grids each begin a group of
// `rows_pad` rows, `stride` elements apart, so the exchange's transformed pair
// grids can land in the padded row layout its contraction reads.  {1, 1, g^3}:
// grids back to back.
struct GridMap {
  int rows, rows_pad;
  long stride;
};

__device__ __forceinline__ long grid_off(const GridMap& m, int b) {
  const int grp = b / m.rows;
  return ((long)grp * m.rows_pad + (b - grp * m.rows)) * m.stride;
}


__device__ __forceinline__ double sfma(double a, double b, double c) { return fma(a, b, c); }
__device__ __forceinline__ float sfma(float a, float b, float c) { return fmaf(a, b, c); }

// complex product with explicit FMAs (no contraction left to the compiler)
template <typename C>
__device__ __forceinline__ C cmul(C a, C b) {
  C r;
  r.x = sfma(a.x, b.x, -(a.y * b.y));
  r.y = sfma(a.x, b.y, a.y * b.x);
  return r;
}
template <typename C>
__device__ __forceinline__ C cadd(C a, C b) {
  C r;
  r.x = a.x + b.x;
  r.y = a.y + b.y;
  return r;
}
template <typename C>
__device__ __forceinline__ C csub(C a, C b) {
  C r;
  r.x = a.x - b.x;
  r.y = a.y - b.y;
  return r;
}
// t1 + i t2 and t1 - i t2
template <typename C>
__device__ __forceinline__ void pm_i(C t1, C t2, C& p, C& m) {
  p.x = t1.x - t2.y;
  p.y = t1.y + t2.x;
  m.x = t1.x + t2.y;
  m.y = t1.y - t2.x;
}
template <typename C> __device__ __forceinline__ C czero();
template <> __device__ __forceinline__ double2 czero<double2>() { return make_double2(0.0, 0.0); }
template <> __device__ __forceinline__ float2 czero<float2>() { return make_float2(0.f, 0.f); }
// twiddle exp(sign 2 pi i m / g), formed in FP64 and rounded once for FP32
template <typename C> __device__ __forceinline__ C twiddle(int m, int g, int sign);
template <> __device__ __forceinline__ double2 twiddle<double2>(int m, int g, int sign) {
  double s, c;
  sincospi((double)(2 * m) / (double)g, &s, &c);
  return make_double2(c, sign < 0 ? -s : s);
}
template <> __device__ __forceinline__ float2 twiddle<float2>(int m, int g, int sign) {
  double s, c;
  sincospi((double)(2 * m) / (double)g, &s, &c);
  return make_float2((float)c, (float)(sign < 0 ? -s : s));
}

// cos and sin of 2 pi m / R for m = 1 .. (R-1)/2, rounded to FP64 from
// extended-precision values
__host__ __device__ constexpr double bf_cos(int R, int m) {
  switch (R * 16 + m) {
    case 3 * 16 + 1: return -0.5;
    case 5 * 16 + 1: return 0.30901699437494745;
    case 5 * 16 + 2: return -0.8090169943749475;
    case 7 * 16 + 1: return 0.6234898018587335;
    case 7 * 16 + 2: return -0.2225209339563144;
    case 7 * 16 + 3: return -0.9009688679024191;
    case 11 * 16 + 1: return 0.8412535328311812;
    case 11 * 16 + 2: return 0.41541501300188644;
    case 11 * 16 + 3: return -0.14231483827328514;
    case 11 * 16 + 4: return -0.6548607339452851;
    case 11 * 16 + 5: return -0.9594929736144974;
    case 13 * 16 + 1: return 0.8854560256532099;
    case 13 * 16 + 2: return 0.5680647467311558;
    case 13 * 16 + 3: return 0.12053668025532305;
    case 13 * 16 + 4: return -0.3546048870425356;
    case 13 * 16 + 5: return -0.7485107481711011;
    case 13 * 16 + 6: return -0.970941817426052;
  }
  return 0.0;
}
__host__ __device__ constexpr double bf_sin(int R, int m) {
  switch (R * 16 + m) {
    case 3 * 16 + 1: return 0.8660254037844386;
    case 5 * 16 + 1: return 0.9510565162951535;
    case 5 * 16 + 2: return 0.5877852522924731;
    case 7 * 16 + 1: return 0.7818314824680298;
    case 7 * 16 + 2: return 0.9749279121818236;
    case 7 * 16 + 3: return 0.4338837391175581;
    case 11 * 16 + 1: return 0.5406408174555976;
    case 11 * 16 + 2: return 0.9096319953545183;
    case 11 * 16 + 3: return 0.9898214418809327;
    case 11 * 16 + 4: return 0.7557495743542583;
    case 11 * 16 + 5: return 0.28173255684142967;
    case 13 * 16 + 1: return 0.46472317204376856;
    case 13 * 16 + 2: return 0.8229838658936564;
    case 13 * 16 + 3: return 0.992708874098054;
    case 13 * 16 + 4: return 0.9350162426854148;
    case 13 * 16 + 5: return 0.6631226582407952;
    case 13 * 16 + 6: return 0.23931566428755777;
  }
  return 0.0;
}
// the same for any m = 1 .. R-1 (cos symmetric, sin antisymmetric about R/2)
__host__ __device__ constexpr double bf_c(int R, int m) {
  return 2 * m < R ? bf_cos(R, m) : bf_cos(R, R - m);
}
__host__ __device__ constexpr double bf_s(int R, int m) {
  return 2 * m < R ? bf_sin(R, m) : -bf_sin(R, R - m);
}

// In-register R-point DFT X_k = sum_r v_r exp(sign 2 pi i r k / R), outputs in
// natural order over the inputs.
template <int R, typename C>
__device__ __forceinline__ void bfly(C (&v)[R], int sign) {
  typedef decltype(v[0].x) T;
  if constexpr (R == 2) {
    const C a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  } else if constexpr (R == 4) {
    const C s02 = cadd(v[0], v[2]), d02 = csub(v[0], v[2]);
    const C s13 = cadd(v[1], v[3]), d13 = csub(v[1], v[3]);
    v[0] = cadd(s02, s13);
    v[2] = csub(s02, s13);
    C p, m;
    pm_i(d02, d13, p, m);
    v[1] = sign > 0 ? p : m;
    v[3] = sign > 0 ? m : p;
  } else {
    constexpr int H = (R - 1) / 2;
    C a[H], b[H];
#pragma unroll
    for (int r = 1; r <= H; ++r) {
      a[r - 1] = cadd(v[r], v[R - r]);
      b[r - 1] = csub(v[r], v[R - r]);
    }
    const C x0 = v[0];
    C s = x0;
#pragma unroll
    for (int r = 0; r < H; ++r) s = cadd(s, a[r]);
#pragma unroll
    for (int k = 1; k <= H; ++k) {
      C t1 = x0, t2 = czero<C>();
#pragma unroll
      for (int r = 1; r <= H; ++r) {
        const T c = (T)bf_c(R, (r * k) % R), sn = (T)bf_s(R, (r * k) % R);
        t1.x = sfma(a[r - 1].x, c, t1.x);
        t1.y = sfma(a[r - 1].y, c, t1.y);
        t2.x = sfma(b[r - 1].x, sn, t2.x);
        t2.y = sfma(b[r - 1].y, sn, t2.y);
      }
      C p, m;
      pm_i(t1, t2, p, m);
      v[k] = sign > 0 ? p : m;
      v[R - k] = sign > 0 ? m : p;
    }
    v[0] = s;
  }
}

// ---------------------------------------------------------------------------
// Passes.  Element (position p, line) of the lpb lines lives at p*ps + line*ls.
// Stage with radix R after Ns (the product of the earlier radices): work item
// (line, j), j < N/R, reads positions j + r N/R, multiplies input r by
// tw[k r N/(Ns R)] (k = j mod Ns; all ones when Ns = 1, skipped: exact), and
// writes output kk to (j/Ns) Ns R + k + kk Ns.
// ---------------------------------------------------------------------------
template <int R, typename C>
__device__ __forceinline__ void radix_item(C (&v)[R], const C* __restrict__ tw, int Ns, int k,
                                           int step, int sign) {
  if (Ns > 1) {
#pragma unroll
    for (int r = 1; r < R; ++r) v[r] = cmul(v[r], tw[k * r * step]);
  }
  bfly<R>(v, sign);
}

template <int R, typename C>
__device__ void pass_radix(const C* __restrict__ in, C* __restrict__ out,
                           const C* __restrict__ tw, int N, int Ns, int lpb, int ps, int ls,
                           int sign) {
  const int NR = N / R;
  const int step = N / (Ns * R);
  for (int it = threadIdx.x; it < NR * lpb; it += blockDim.x) {
    const int line = it % lpb, j = it / lpb;
    const int k = j % Ns;
    C v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) v[r] = in[(j + r * NR) * ps + line * ls];
    radix_item<R>(v, tw, Ns, k, step, sign);
    const int idx = (j / Ns) * Ns * R + k;
#pragma unroll
    for (int kk = 0; kk < R; ++kk) out[(idx + kk * Ns) * ps + line * ls] = v[kk];
  }
}

// In-place form: at most ip_items(R, MI) work items per thread, all inputs read
// before the block synchronises (the host checks that the items fit).
__host__ __device__ constexpr int ip_items(int R, int MI) { return R >= 11 ? 1 : MI; }

template <int R, int MI, typename C>
__device__ void pass_radix_ip(C* __restrict__ buf, const C* __restrict__ tw, int N, int Ns,
                              int lpb, int ps, int ls, int sign) {
  constexpr int M = ip_items(R, MI);
  const int NR = N / R;
  const int step = N / (Ns * R);
  const int total = NR * lpb;
  C v[M][R];
#pragma unroll
  for (int m = 0; m < M; ++m) {
    const int it = threadIdx.x + m * blockDim.x;
    if (it < total) {
      const int line = it % lpb, j = it / lpb;
#pragma unroll
      for (int r = 0; r < R; ++r) v[m][r] = buf[(j + r * NR) * ps + line * ls];
    }
  }
  __syncthreads();
#pragma unroll
  for (int m = 0; m < M; ++m) {
    const int it = threadIdx.x + m * blockDim.x;
    if (it < total) {
      const int line = it % lpb, j = it / lpb;
      const int k = j % Ns;
      radix_item<R>(v[m], tw, Ns, k, step, sign);
      const int idx = (j / Ns) * Ns * R + k;
#pragma unroll
      for (int kk = 0; kk < R; ++kk) buf[(idx + kk * Ns) * ps + line * ls] = v[m][kk];
    }
  }
  __syncthreads();
}

// Generic odd factor R, sub-pass 1 (in place, one thread per pair): twiddle
// inputs r and R-r of every group and store (a_r, b_r) over them.
template <typename C>
__device__ void generic_pairs(C* __restrict__ buf, const C* __restrict__ tw, int N, int Ns, int R,
                              int lpb, int ps, int ls) {
  const int NR = N / R, H = (R - 1) / 2;
  const int step = N / (Ns * R);
  for (int it = threadIdx.x; it < NR * lpb * H; it += blockDim.x) {
    const int line = it % lpb, q = it / lpb;
    const int j = q % NR, r = 1 + q / NR;
    const int k = j % Ns;
    C* px = buf + (j + r * NR) * ps + line * ls;
    C* py = buf + (j + (R - r) * NR) * ps + line * ls;
    C x = *px, y = *py;
    if (Ns > 1) {
      x = cmul(x, tw[k * r * step]);
      y = cmul(y, tw[k * (R - r) * step]);
    }
    *px = cadd(x, y);
    *py = csub(x, y);
  }
}

// Sub-pass 2, the sums of one work item (line, j, block of FFT_PAIRS pairs):
// t1 = v_0 + sum_r a_r cos, t2 = sum_r b_r s sin for each kk of the block (kk
// past (R-1)/2 marked 0, computed on index 0 and not stored), s0 = X_0.
template <typename C>
__device__ __forceinline__ void generic_sums(const C* __restrict__ in, const C* __restrict__ tw,
                                             int R, int NR, int lpb, int ps, int ls, int it,
                                             C (&t1)[FFT_PAIRS], C (&t2)[FFT_PAIRS], C& s0,
                                             int (&kk)[FFT_PAIRS]) {
  const int H = (R - 1) / 2;
  const int line = it % lpb, q = it / lpb;
  const int j = q % NR, blk = q / NR;
  const C* base = in + j * ps + line * ls;
  const C v0 = base[0];
  int m[FFT_PAIRS];
#pragma unroll
  for (int p = 0; p < FFT_PAIRS; ++p) {
    kk[p] = 1 + blk * FFT_PAIRS + p;
    if (kk[p] > H) kk[p] = 0;
    m[p] = 0;
    t1[p] = v0;
    t2[p] = czero<C>();
  }
  s0 = v0;
  const int rs = NR * ps;
  for (int r = 1; r <= H; ++r) {
    const C a = base[r * rs], b = base[(R - r) * rs];
    s0 = cadd(s0, a);
#pragma unroll
    for (int p = 0; p < FFT_PAIRS; ++p) {
      m[p] += kk[p];
      if (m[p] >= R) m[p] -= R;
      const C w = tw[m[p] * NR];
      t1[p].x = sfma(a.x, w.x, t1[p].x);
      t1[p].y = sfma(a.y, w.x, t1[p].y);
      t2[p].x = sfma(b.x, w.y, t2[p].x);
      t2[p].y = sfma(b.y, w.y, t2[p].y);
    }
  }
}

template <typename C>
__device__ __forceinline__ void generic_store(C* __restrict__ out, int R, int NR, int Ns, int lpb,
                                              int ps, int ls, int it, const C (&t1)[FFT_PAIRS],
                                              const C (&t2)[FFT_PAIRS], C s0,
                                              const int (&kk)[FFT_PAIRS]) {
  const int line = it % lpb, q = it / lpb;
  const int j = q % NR, blk = q / NR;
  const int k = j % Ns;
  const int idx = (j / Ns) * Ns * R + k;
  C* o = out + line * ls;
#pragma unroll
  for (int p = 0; p < FFT_PAIRS; ++p) {
    if (kk[p]) {
      C xp, xm;
      pm_i(t1[p], t2[p], xp, xm);
      o[(idx + kk[p] * Ns) * ps] = xp;
      o[(idx + (R - kk[p]) * Ns) * ps] = xm;
    }
  }
  if (blk == 0) o[idx * ps] = s0;
}

template <typename C>
__device__ void pass_generic(C* __restrict__ in, C* __restrict__ out, const C* __restrict__ tw,
                             int N, int Ns, int R, int lpb, int ps, int ls) {
  generic_pairs(in, tw, N, Ns, R, lpb, ps, ls);
  __syncthreads();
  const int NR = N / R;
  const int nblk = ((R - 1) / 2 + FFT_PAIRS - 1) / FFT_PAIRS;
  for (int it = threadIdx.x; it < NR * lpb * nblk; it += blockDim.x) {
    C t1[FFT_PAIRS], t2[FFT_PAIRS], s0;
    int kk[FFT_PAIRS];
    generic_sums(in, tw, R, NR, lpb, ps, ls, it, t1, t2, s0, kk);
    generic_store(out, R, NR, Ns, lpb, ps, ls, it, t1, t2, s0, kk);
  }
}

template <int MI, typename C>
__device__ void pass_generic_ip(C* __restrict__ buf, const C* __restrict__ tw, int N, int Ns,
                                int R, int lpb, int ps, int ls) {
  generic_pairs(buf, tw, N, Ns, R, lpb, ps, ls);
  __syncthreads();
  const int NR = N / R;
  const int nblk = ((R - 1) / 2 + FFT_PAIRS - 1) / FFT_PAIRS;
  const int total = NR * lpb * nblk;
  C t1[MI][FFT_PAIRS], t2[MI][FFT_PAIRS], s0[MI];
  int kk[MI][FFT_PAIRS];
#pragma unroll
  for (int m = 0; m < MI; ++m) {
    const int it = threadIdx.x + m * blockDim.x;
    if (it < total) generic_sums(buf, tw, R, NR, lpb, ps, ls, it, t1[m], t2[m], s0[m], kk[m]);
  }
  __syncthreads();
#pragma unroll
  for (int m = 0; m < MI; ++m) {
    const int it = threadIdx.x + m * blockDim.x;
    if (it < total) generic_store(buf, R, NR, Ns, lpb, ps, ls, it, t1[m], t2[m], s0[m], kk[m]);
  }
  __syncthreads();
}

// All passes of one transform of length N over lpb lines between two buffers;
// returns the buffer that holds the result (in or out).  `in` is overwritten.
template <typename C>
__device__ C* run_passes(C* in, C* out, const C* __restrict__ tw, const FftPlan& plan, int N,
                         int lpb, int ps, int ls, int sign) {
  int Ns = 1;
  for (int s = 0; s < plan.nst; ++s) {
    const int R = plan.rad[s];
    switch (R) {
      case 2: pass_radix<2>(in, out, tw, N, Ns, lpb, ps, ls, sign); break;
      case 3: pass_radix<3>(in, out, tw, N, Ns, lpb, ps, ls, sign); break;
      case 4: pass_radix<4>(in, out, tw, N, Ns, lpb, ps, ls, sign); break;
      case 5: pass_radix<5>(in, out, tw, N, Ns, lpb, ps, ls, sign); break;
      case 7: pass_radix<7>(in, out, tw, N, Ns, lpb, ps, ls, sign); break;
      case 11: pass_radix<11>(in, out, tw, N, Ns, lpb, ps, ls, sign); break;
      case 13: pass_radix<13>(in, out, tw, N, Ns, lpb, ps, ls, sign); break;
      default: pass_generic(in, out, tw, N, Ns, R, lpb, ps, ls); break;
    }
    __syncthreads();
    C* t = in;
    in = out;
    out = t;
    Ns *= R;
  }
  return in;
}

template <int MI, typename C>
__device__ void run_passes_ip(C* buf, const C* __restrict__ tw, const FftPlan& plan, int N,
                              int lpb, int ps, int ls, int sign) {
  int Ns = 1;
  for (int s = 0; s < plan.nst; ++s) {
    const int R = plan.rad[s];
    switch (R) {
      case 2: pass_radix_ip<2, MI>(buf, tw, N, Ns, lpb, ps, ls, sign); break;
      case 3: pass_radix_ip<3, MI>(buf, tw, N, Ns, lpb, ps, ls, sign); break;
      case 4: pass_radix_ip<4, MI>(buf, tw, N, Ns, lpb, ps, ls, sign); break;
      case 5: pass_radix_ip<5, MI>(buf, tw, N, Ns, lpb, ps, ls, sign); break;
      case 7: pass_radix_ip<7, MI>(buf, tw, N, Ns, lpb, ps, ls, sign); break;
      case 11: pass_radix_ip<11, MI>(buf, tw, N, Ns, lpb, ps, ls, sign); break;
      case 13: pass_radix_ip<13, MI>(buf, tw, N, Ns, lpb, ps, ls, sign); break;
      default: pass_generic_ip<MI>(buf, tw, N, Ns, R, lpb, ps, ls); break;
    }
    Ns *= R;
  }
}

static bool fft_radix_known(int R) { return R <= 5 || R == 7 || R == 11 || R == 13; }

// Whether every pass of a length-N transform over lpb lines fits the in-place
// passes with at most MI work items per thread and nt threads.
static bool fft_inplace_fits(const FftPlan& plan, int N, int lpb, int nt, int MI) {
  for (int s = 0; s < plan.nst; ++s) {
    const int R = plan.rad[s];
    const long items = fft_radix_known(R)
                           ? (long)(N / R) * lpb
                           : (long)(N / R) * lpb * (((R - 1) / 2 + FFT_PAIRS - 1) / FFT_PAIRS);
    const int per = fft_radix_known(R) ? ip_items(R, MI) : MI;
    if (items > (long)per * nt) return false;
  }
  return true;
}

template <typename C, int MI>
__global__ __launch_bounds__(FFT_THREADS) void k_fft_axis(C* __restrict__ data, int g, int axis,
                                                          int lpb, FftPlan plan, int sign,
                                                          GridMap map) {
  extern __shared__ unsigned char sm_raw[];
  C* bufA = reinterpret_cast<C*>(sm_raw);
  C* bufB = bufA + g * lpb;
  C* tw = MI ? bufB : bufB + g * lpb;
  const long gg = (long)g * g;
  C* base = data + grid_off(map, blockIdx.y);
  const int line0 = blockIdx.x * lpb;
  for (int m = threadIdx.x; m < g; m += blockDim.x) tw[m] = twiddle<C>(m, g, sign);
  const long stride = axis == 2 ? 1 : (axis == 1 ? g : gg);
  for (int e = threadIdx.x; e < g * lpb; e += blockDim.x) {
    int p, line;
    if (axis == 2) {
      line = e / g;
      p = e - line * g;
    } else {
      line = e % lpb;
      p = e / lpb;
    }
    const long l = line0 + line;
    C val = czero<C>();
    if (l < gg) {
      const long lb = axis == 2 ? l * g : (axis == 1 ? (l / g) * gg + (l % g) : l);
      val = base[lb + p * stride];
    }
    bufA[p * lpb + line] = val;
  }
  __syncthreads();
  C* in = bufA;
  if (MI)
    run_passes_ip<MI ? MI : 1>(bufA, tw, plan, g, lpb, lpb, 1, sign);
  else
    in = run_passes(bufA, bufB, tw, plan, g, lpb, lpb, 1, sign);
  for (int e = threadIdx.x; e < g * lpb; e += blockDim.x) {
    int p, line;
    if (axis == 2) {
      line = e / g;
      p = e - line * g;
    } else {
      line = e % lpb;
      p = e / lpb;
    }
    const long l = line0 + line;
    if (l < gg) {
      const long lb = axis == 2 ? l * g : (axis == 1 ? (l / g) * gg + (l % g) : l);
      base[lb + p * stride] = in[p * lpb + line];
    }
  }
}

// The two contiguous axes in one launch: a workgroup loads the whole (y, z)
// plane of one x (g^2 contiguous points, coalesced), transforms along z
// (lines y: element (z, y) at y*g + z) and then along y (lines z: element
// (y, z) at y*g + z) in LDS, and writes the plane back.  Together with the
// strided x pass this reads and writes the grid twice instead of three times.
// NT threads (256, 512 or 1024): each output is formed the same way whichever
// thread forms it, so the thread count and the in-place form change the
// schedule only.
#define FFT_PLANE_MAX_LDS (156 * 1024)

// The exchange's pair densities formed on the plane pass's loads: grid m * nao + p
// of the batch is phi_p (psi_i + i psi_j) of table row m = (geometry, i, j) (j < 0:
// none), the product hqc_pair_pack stores (an FP64 multiply, rounded once to FP32
// for complex64), so the pack's write of the grids and this pass's read of them
// are saved.  Unused (null) for a plain transform.
struct PairSrc {
  const double* phi;   // (geometries, nao_pad, ld) AO values
  const double* psi;   // (geometries, nmo, ld) weighted orbitals
  const int* table;    // (mp, 3)
  int nao, nao_pad, nmo;
  long ld;
};

__device__ __forceinline__ double2 pair_c(double a, double b, double2*) { return make_double2(a, b); }
__device__ __forceinline__ float2 pair_c(double a, double b, float2*) {
  return make_float2((float)a, (float)b);
}

template <int NT, typename C, int MI, bool PACK>
__global__ __launch_bounds__(NT) void k_fft_plane(C* __restrict__ data, int g, FftPlan plan,
                                                  int sign, PairSrc src, GridMap map) {
  extern __shared__ unsigned char sm_raw[];
  const int gg = g * g;
  C* bufA = reinterpret_cast<C*>(sm_raw);
  C* bufB = bufA + gg;
  C* tw = MI ? bufB : bufB + gg;
  C* plane = data + grid_off(map, blockIdx.y) + (long)blockIdx.x * gg;
  for (int m = threadIdx.x; m < g; m += blockDim.x) tw[m] = twiddle<C>(m, g, sign);
  if (PACK) {
    const int m = blockIdx.y / src.nao, p = blockIdx.y - m * src.nao;
    const int b = src.table[3 * m], i = src.table[3 * m + 1], j = src.table[3 * m + 2];
    const long off = (long)blockIdx.x * gg;
    const double* F = src.phi + ((long)b * src.nao_pad + p) * src.ld + off;
    const double* A = src.psi + ((long)b * src.nmo + i) * src.ld + off;
    const double* B = src.psi + ((long)b * src.nmo + (j >= 0 ? j : i)) * src.ld + off;
    for (int e = threadIdx.x; e < gg; e += blockDim.x) {
      const double f = F[e];
      const double c = j >= 0 ? B[e] : 0.0;
      bufA[e] = pair_c(f * A[e], f * c, (C*)nullptr);
    }
  } else {
    for (int e = threadIdx.x; e < gg; e += blockDim.x) bufA[e] = plane[e];
  }
  __syncthreads();
  if (MI) {
    run_passes_ip<MI ? MI : 1>(bufA, tw, plan, g, g, 1, g, sign);      // along z
    run_passes_ip<MI ? MI : 1>(bufA, tw, plan, g, g, g, 1, sign);      // along y
    for (int e = threadIdx.x; e < gg; e += blockDim.x) plane[e] = bufA[e];
    return;
  }
  C* r = run_passes(bufA, bufB, tw, plan, g, g, 1, g, sign);          // along z
  C* o = (r == bufA) ? bufB : bufA;
  r = run_passes(r, o, tw, plan, g, g, g, 1, sign);                    // along y
  for (int e = threadIdx.x; e < gg; e += blockDim.x) plane[e] = r[e];
}

static int make_plan(int g, FftPlan& plan) {
  plan.nst = 0;
  int N = g;
  const int pref[] = {4, 2, 3, 5, 7, 11, 13};
  for (int r : pref) {
    while (N % r == 0) {
      if (plan.nst >= FFT_MAX_STAGES) return HQC_E_SIZE;
      plan.rad[plan.nst++] = r;
      N /= r;
    }
  }
  for (int r = 17; N > 1; r += 2) {
    while (N % r == 0) {
      if (plan.nst >= FFT_MAX_STAGES) return HQC_E_SIZE;
      plan.rad[plan.nst++] = r;
      N /= r;
    }
  }
  return HQC_OK;
}

template <int NT, typename C, bool PACK>
static const void* plane_kernel(int mi) {
  return mi == 0 ? (const void*)k_fft_plane<NT, C, 0, PACK>
                 : (mi == 1 ? (const void*)k_fft_plane<NT, C, 1, PACK>
                            : (const void*)k_fft_plane<NT, C, 2, PACK>);
}

template <int NT, typename C, bool PACK>
static void launch_plane(int mi, dim3 grid, size_t lds, hipStream_t st, C* data, int g,
                         const FftPlan& plan, int sign, const PairSrc& src, const GridMap& map) {
  if (mi == 0)
    hipLaunchKernelGGL((k_fft_plane<NT, C, 0, PACK>), grid, dim3(NT), lds, st, data, g, plan, sign,
                       src, map);
  else if (mi == 1)
    hipLaunchKernelGGL((k_fft_plane<NT, C, 1, PACK>), grid, dim3(NT), lds, st, data, g, plan, sign,
                       src, map);
  else
    hipLaunchKernelGGL((k_fft_plane<NT, C, 2, PACK>), grid, dim3(NT), lds, st, data, g, plan, sign,
                       src, map);
}

// Whether the in-place passes are wanted by default: in FP64 only, and only where
// two buffers would keep a CU below 16 waves (more than 40 KB of LDS per 256
// threads).  Measured on 21-304 grids: in place 49^3 0.30 -> 0.24 ms, 61^3 0.48 ->
// 0.34 ms at 256 threads, 55^3 0.161 -> 0.144 ms at 512; two buffers faster where
// LDS does not limit (29^3 at 256, 45^3 at 512: fewer registers and barriers) and
// in FP32 (the same occupancy either way).
static bool inplace_wanted(int g, int lpb, int nt, size_t csize) {
  return csize >= 16 && (size_t)2 * g * lpb * csize > (size_t)40 * 1024 * (nt / 256);
}

// Work items per thread of the in-place passes for a transform over lpb lines
// with nt threads (0: the two-buffer passes): tuned values when they fit, else by
// default 2 items per thread where they fit, else 1, where the in-place form is
// wanted.
static int inplace_items(const FftPlan& plan, int g, int lpb, int nt, size_t csize) {
  if (hqc_tuned(HQC_TUNE_FFT_TWO_BUFFER, 0)) return 0;
  const int want = hqc_tuned(HQC_TUNE_FFT_IP_ITEMS, 0);
  if (want) return fft_inplace_fits(plan, g, lpb, nt, want) ? want : 0;
  if (!inplace_wanted(g, lpb, nt, csize)) return 0;
  for (int mi = 2; mi >= 1; --mi)
    if (fft_inplace_fits(plan, g, lpb, nt, mi)) return mi;
  return 0;
}

// pack (non-null): the batch's grids are formed from pack's pair source on the plane
// pass (which must then be taken: HQC_E_SIZE where the plane does not fit LDS)
template <typename C>
static int fft3d(C* data, int B, int g, int sign, void* stream, const PairSrc* pack = nullptr,
                 const GridMap* gmap = nullptr) {
  const GridMap map = gmap ? *gmap : GridMap{1, 1, (long)g * g * g};
  if (B < 0 || g <= 0 || (sign != 1 && sign != -1)) return HQC_E_ARG;
  if (g > FFT_MAX_G) return HQC_E_SIZE;
  if (B == 0) return HQC_OK;
  if (g == 1 && !pack) return HQC_OK;
  FftPlan plan;
  if (make_plan(g, plan) != HQC_OK) return HQC_E_SIZE;
  hipStream_t st = (hipStream_t)stream;
  const long lines = (long)g * g;
  // threads per plane: with enough planes to fill the GPU, 256 (several planes
  // per CU) up to 61^2 in FP64 and 49^2 in FP32, 512 above; 1024 for few planes
  int nt = hqc_tuned(HQC_TUNE_FFT_PLANE_THREADS, 0);
  if (!nt) {
    const bool f64 = sizeof(C) == 16;
    if ((long)g * B < 512) {
      nt = 1024;
    } else if (g <= (f64 ? 61 : 49)) {
      nt = 256;
      // FP64 planes that want the in-place form but whose passes do not fit 256
      // threads (45^2 .. 57^2: radix-3/5 passes of 600-1100 items) take 512
      // threads, in place or with two buffers as inplace_items decides there
      // (45, 51, 55, 57: 0.155 -> 0.140, 0.225 -> 0.197, 0.179 -> 0.144, 0.276 ->
      // 0.234 ms per 32-65 grids)
      if (f64 && inplace_wanted(g, g, 256, sizeof(C)) && !fft_inplace_fits(plan, g, g, 256, 1) &&
          !fft_inplace_fits(plan, g, g, 256, 2))
        nt = 512;
    } else {
      nt = 512;
    }
  }
  const int mi = inplace_items(plan, g, g, nt, sizeof(C));
  const size_t plane_lds = (size_t)((mi ? 1 : 2) * g * g + g) * sizeof(C);
  int first_axis = 2;
  if (pack && (plane_lds > FFT_PLANE_MAX_LDS || B > 65535)) return HQC_E_SIZE;
  if (plane_lds <= FFT_PLANE_MAX_LDS) {
    const PairSrc src = pack ? *pack : PairSrc{nullptr, nullptr, nullptr, 0, 0, 0, 0};
    const void* kf =
        pack ? (nt == 256 ? plane_kernel<256, C, true>(mi)
                          : (nt == 512 ? plane_kernel<512, C, true>(mi) : plane_kernel<1024, C, true>(mi)))
             : (nt == 256 ? plane_kernel<256, C, false>(mi)
                          : (nt == 512 ? plane_kernel<512, C, false>(mi) : plane_kernel<1024, C, false>(mi)));
    if (plane_lds > 64 * 1024 &&
        hipFuncSetAttribute(kf, hipFuncAttributeMaxDynamicSharedMemorySize, (int)plane_lds) !=
            hipSuccess)
      (void)hipGetLastError();
    const dim3 grid(g, B);
    if (pack) {
      if (nt == 256)
        launch_plane<256, C, true>(mi, grid, plane_lds, st, data, g, plan, sign, src, map);
      else if (nt == 512)
        launch_plane<512, C, true>(mi, grid, plane_lds, st, data, g, plan, sign, src, map);
      else
        launch_plane<1024, C, true>(mi, grid, plane_lds, st, data, g, plan, sign, src, map);
    } else {
      if (nt == 256)
        launch_plane<256, C, false>(mi, grid, plane_lds, st, data, g, plan, sign, src, map);
      else if (nt == 512)
        launch_plane<512, C, false>(mi, grid, plane_lds, st, data, g, plan, sign, src, map);
      else
        launch_plane<1024, C, false>(mi, grid, plane_lds, st, data, g, plan, sign, src, map);
    }
    HQC_CHECK_LAUNCH();
    first_axis = 0;
  }
  if (g == 1) return HQC_OK;
  // remaining axes line by line; more lines per block on the strided x axis
  for (int axis = first_axis; axis >= 0; --axis) {
    // lines per block: 32 on the strided x axis (longer coalesced segments), 16
    // otherwise, and never more than 64 KB of LDS for two buffers
    const int fit = (int)((64 * 1024 / sizeof(C) - g) / (2 * g));
    const int lpb = max(1, min(axis == 0 ? 32 : 16, fit));
    const int ami = inplace_items(plan, g, lpb, FFT_THREADS, sizeof(C));
    const size_t lds = (size_t)((ami ? 1 : 2) * g * lpb + g) * sizeof(C);
    const dim3 grid(hqc_div_up(lines, lpb), B);
    if (ami == 0)
      hipLaunchKernelGGL((k_fft_axis<C, 0>), grid, dim3(FFT_THREADS), lds, st, data, g, axis,
                         lpb, plan, sign, map);
    else if (ami == 1)
      hipLaunchKernelGGL((k_fft_axis<C, 1>), grid, dim3(FFT_THREADS), lds, st, data, g, axis,
                         lpb, plan, sign, map);
    else
      hipLaunchKernelGGL((k_fft_axis<C, 2>), grid, dim3(FFT_THREADS), lds, st, data, g, axis,
                         lpb, plan, sign, map);
    HQC_CHECK_LAUNCH();
  }
  return HQC_OK;
}

extern "C" int hqc_fft3d_c2c_f64(double* data, int B, int g, int sign, void* stream) {
  return fft3d<double2>(reinterpret_cast<double2*>(data), B, g, sign, stream);
}

extern "C" int hqc_fft3d_c2c_f32(float* data, int B, int g, int sign, void* stream) {
  return fft3d<float2>(reinterpret_cast<float2*>(data), B, g, sign, stream);
}

template <typename C>
static int pair_pack_fft(const double* phi, const double* psi, const int* table, int mp, int nao,
                         int nao_pad, int nmo, int g, int ld, C* z, int z_rows, int z_stride,
                         int sign, void* stream) {
  if (mp < 0 || nao <= 0 || nao > nao_pad || nmo <= 0 || g <= 0 || z_rows < nao) return HQC_E_ARG;
  const long N = (long)g * g * g;
  if ((long)ld < N || (long)z_stride < N || (long)mp * nao > 65535) return HQC_E_SIZE;
  if (mp == 0) return HQC_OK;
  if (!phi || !psi || !table || !z) return HQC_E_ARG;
  const PairSrc src{phi, psi, table, nao, nao_pad, nmo, (long)ld};
  const GridMap map{nao, z_rows, (long)z_stride};
  return fft3d<C>(z, mp * nao, g, sign, stream, &src, &map);
}

extern "C" int hqc_pair_pack_fft(const double* phi, const double* psi, const int* table, int mp,
                                 int nao, int nao_pad, int nmo, int g, int ld, double* z, int z_rows,
                                 int z_stride, int sign, void* stream) {
  return pair_pack_fft<double2>(phi, psi, table, mp, nao, nao_pad, nmo, g, ld,
                                reinterpret_cast<double2*>(z), z_rows, z_stride, sign, stream);
}

extern "C" int hqc_pair_pack_fft_c64(const double* phi, const double* psi, const int* table,
                                     int mp, int nao, int nao_pad, int nmo, int g, int ld, float* z,
                                     int z_rows, int z_stride, int sign, void* stream) {
  return pair_pack_fft<float2>(phi, psi, table, mp, nao, nao_pad, nmo, g, ld,
                               reinterpret_cast<float2*>(z), z_rows, z_stride, sign, stream);
}
