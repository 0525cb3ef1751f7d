// Batched 3-D complex FFT for the odd (and even) grids of the GPW solver.
//
// Up to g = 69 the y and z axes are done together on whole planes in LDS
// (k_fft_plane), then the x axis by lines; larger grids go axis by axis.  A
// line block stages `LPB` lines of length g in LDS
// (position-major, line-minor so that strided axes load coalesced), runs the
// mixed-radix Stockham passes between two LDS buffers, and writes the lines
// back in place.  Radix butterflies are specialised for 2,3,4,5,7,11,13; any
// other prime factor uses a direct per-output DFT pass.  Twiddles come from a
// length-g table exp(sign 2 pi i m/g) built in LDS with sincospi.
#include "common.h"

#define FFT_MAX_G 128
#define FFT_MAX_STAGES 12
#define FFT_THREADS 256

struct FftPlan {
  int nst;
  int rad[FFT_MAX_STAGES];
};

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
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

// Stockham pass with compile-time radix R (see the header comment).  Element
// (position p, line) of the lpb lines lives at p*ps + line*ls in LDS.
template <int R, typename C>
__device__ void pass_r(const C* __restrict__ in, C* __restrict__ out,
                       const C* __restrict__ tw, int N, int Ns, int lpb, int ps, int ls) {
  const int NR = N / R;
  const int step = N / (Ns * R);
  for (int it = threadIdx.x; it < NR * lpb; it += blockDim.x) {
    const int line = it % lpb, j = it / lpb;
    const int k = j % Ns;
    C v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) v[r] = in[(j + r * NR) * ps + line * ls];
#pragma unroll
    for (int r = 1; r < R; ++r) v[r] = cmul(v[r], tw[k * r * step]);
    const int idx = (j / Ns) * Ns * R + k;
#pragma unroll
    for (int kk = 0; kk < R; ++kk) {
      C acc = v[0];
#pragma unroll
      for (int r = 1; r < R; ++r) {
        const C wr = tw[((r * kk) % R) * NR];
        acc.x += v[r].x * wr.x - v[r].y * wr.y;
        acc.y += v[r].x * wr.y + v[r].y * wr.x;
      }
      out[(idx + kk * Ns) * ps + line * ls] = acc;
    }
  }
}

// Generic radix: one work item per output element.
template <typename C>
__device__ void pass_generic(const C* __restrict__ in, C* __restrict__ out,
                             const C* __restrict__ tw, int N, int Ns, int R, int lpb,
                             int ps, int ls) {
  const int NR = N / R;
  const int step = N / (Ns * R);
  for (int it = threadIdx.x; it < N * lpb; it += blockDim.x) {
    const int line = it % lpb;
    const int q = it / lpb;  // q = j*R + kk
    const int j = q / R, kk = q % R;
    const int k = j % Ns;
    C acc = czero<C>();
    for (int r = 0; r < R; ++r) {
      C x = in[(j + r * NR) * ps + line * ls];
      if (r) x = cmul(x, tw[k * r * step]);
      const C w = tw[((r * kk) % R) * NR];
      acc.x += x.x * w.x - x.y * w.y;
      acc.y += x.x * w.y + x.y * w.x;
    }
    const int idx = (j / Ns) * Ns * R + k;
    out[(idx + kk * Ns) * ps + line * ls] = acc;
  }
}

// A whole transform of prime length N > 13 in one pass (direct DFT, the only
// stage).  Work item = (line, block of PR outputs): the inputs of the line are
// read once each and accumulated into PR outputs; the twiddle index of output
// q after input r is (r * q) mod N, stepped by q with one conditional subtract.
#define FFT_PRIME_PR 8

template <typename C>
__device__ void pass_prime(const C* __restrict__ in, C* __restrict__ out,
                           const C* __restrict__ tw, int N, int lpb, int ps, int ls) {
  const int nblk = (N + FFT_PRIME_PR - 1) / FFT_PRIME_PR;
  for (int it = threadIdx.x; it < nblk * lpb; it += blockDim.x) {
    const int line = it % lpb, q0 = (it / lpb) * FFT_PRIME_PR;
    C acc[FFT_PRIME_PR];
    int idx[FFT_PRIME_PR];
#pragma unroll
    for (int j = 0; j < FFT_PRIME_PR; ++j) {
      acc[j] = czero<C>();
      idx[j] = 0;
    }
    for (int r = 0; r < N; ++r) {
      const C x = in[r * ps + line * ls];
#pragma unroll
      for (int j = 0; j < FFT_PRIME_PR; ++j) {
        const C w = tw[idx[j]];
        acc[j].x += x.x * w.x - x.y * w.y;
        acc[j].y += x.x * w.y + x.y * w.x;
        // (r + 1) q mod N for q = q0 + j; outputs past N (discarded) keep q = 0
        const int q = q0 + j < N ? q0 + j : 0;
        const int t = idx[j] + q;
        idx[j] = t >= N ? t - N : t;
      }
    }
#pragma unroll
    for (int j = 0; j < FFT_PRIME_PR; ++j)
      if (q0 + j < N) out[(q0 + j) * ps + line * ls] = acc[j];
  }
}

// All Stockham passes of one transform of length N over lpb lines; returns the
// buffer that holds the result (in or out).
template <typename C>
__device__ C* run_passes(C* in, C* out, const C* __restrict__ tw, const FftPlan& plan, int N,
                         int lpb, int ps, int ls) {
  if (plan.nst == 1 && plan.rad[0] > 13) {
    pass_prime(in, out, tw, N, lpb, ps, ls);
    __syncthreads();
    return out;
  }
  int Ns = 1;
  for (int s = 0; s < plan.nst; ++s) {
    const int R = plan.rad[s];
    switch (R) {
      case 2: pass_r<2>(in, out, tw, N, Ns, lpb, ps, ls); break;
      case 3: pass_r<3>(in, out, tw, N, Ns, lpb, ps, ls); break;
      case 4: pass_r<4>(in, out, tw, N, Ns, lpb, ps, ls); break;
      case 5: pass_r<5>(in, out, tw, N, Ns, lpb, ps, ls); break;
      case 7: pass_r<7>(in, out, tw, N, Ns, lpb, ps, ls); break;
      case 11: pass_r<11>(in, out, tw, N, Ns, lpb, ps, ls); break;
      case 13: pass_r<13>(in, out, tw, N, Ns, lpb, ps, ls); break;
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

template <typename C>
__global__ __launch_bounds__(FFT_THREADS) void k_fft_axis(C* __restrict__ data, int g, int axis,
                                                          int lpb, FftPlan plan, int sign) {
  extern __shared__ unsigned char sm_raw[];
  C* bufA = reinterpret_cast<C*>(sm_raw);
  C* bufB = bufA + g * lpb;
  C* tw = bufB + g * lpb;
  const long gg = (long)g * g;
  C* base = data + (long)blockIdx.y * gg * g;
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
  C* in = run_passes(bufA, bufB, tw, plan, g, lpb, lpb, 1);
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
// NT threads: 1024, or 256 for planes up to 51^2 when there are at least two
// planes per CU (the radix passes of a 49^2 plane have 343 to 1225 work items;
// 128 grids of 49^3: whole transform 0.452 -> 0.393 ms, the same results bit
// for bit, since each output is formed the same way whichever thread forms it;
// 65^3 and a few planes per CU run slower at 256).
#define FFT_PLANE_MAX_LDS (156 * 1024)

template <int NT, typename C>
__global__ __launch_bounds__(NT) void k_fft_plane(C* __restrict__ data, int g, FftPlan plan,
                                                  int sign) {
  extern __shared__ unsigned char sm_raw[];
  const int gg = g * g;
  C* bufA = reinterpret_cast<C*>(sm_raw);
  C* bufB = bufA + gg;
  C* tw = bufB + gg;
  C* plane = data + ((long)blockIdx.y * g + blockIdx.x) * gg;
  for (int m = threadIdx.x; m < g; m += blockDim.x) tw[m] = twiddle<C>(m, g, sign);
  for (int e = threadIdx.x; e < gg; e += blockDim.x) bufA[e] = plane[e];
  __syncthreads();
  C* r = run_passes(bufA, bufB, tw, plan, g, g, 1, g);                // along z
  C* o = (r == bufA) ? bufB : bufA;
  r = run_passes(r, o, tw, plan, g, g, g, 1);                          // along y
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

template <typename C>
static int fft3d(C* data, int B, int g, int sign, void* stream) {
  if (B < 0 || g <= 0 || (sign != 1 && sign != -1)) return HQC_E_ARG;
  if (g > FFT_MAX_G) return HQC_E_SIZE;
  if (B == 0 || g == 1) return HQC_OK;
  FftPlan plan;
  if (make_plan(g, plan) != HQC_OK) return HQC_E_SIZE;
  const long lines = (long)g * g;
  const size_t plane_lds = (size_t)(2 * g * g + g) * sizeof(C);
  int first_axis = 2;
  if (plane_lds <= FFT_PLANE_MAX_LDS) {
    const int nt = hqc_tuned(HQC_TUNE_FFT_PLANE_THREADS, 0);
    const bool small = nt ? nt == 256 : (g <= 51 && (long)g * B >= 512);
    const void* kf = small ? (const void*)k_fft_plane<256, C> : (const void*)k_fft_plane<1024, C>;
    if (plane_lds > 64 * 1024 &&
        hipFuncSetAttribute(kf, hipFuncAttributeMaxDynamicSharedMemorySize, (int)plane_lds) !=
            hipSuccess)
      (void)hipGetLastError();
    if (small)
      hipLaunchKernelGGL((k_fft_plane<256, C>), dim3(g, B), dim3(256), plane_lds,
                         (hipStream_t)stream, data, g, plan, sign);
    else
      hipLaunchKernelGGL((k_fft_plane<1024, C>), dim3(g, B), dim3(1024), plane_lds,
                         (hipStream_t)stream, data, g, plan, sign);
    HQC_CHECK_LAUNCH();
    first_axis = 0;
  }
  // remaining axes line by line; more lines per block on the strided x axis
  for (int axis = first_axis; axis >= 0; --axis) {
    // lines per block: 32 on the strided x axis (longer coalesced segments), 16
    // otherwise, and never more than 64 KB of LDS
    const int fit = (int)((64 * 1024 / sizeof(C) - g) / (2 * g));
    const int lpb = max(1, min(axis == 0 ? 32 : 16, fit));
    const size_t lds = (size_t)(2 * g * lpb + g) * sizeof(C);
    dim3 grid(hqc_div_up(lines, lpb), B);
    hipLaunchKernelGGL((k_fft_axis<C>), grid, dim3(FFT_THREADS), lds, (hipStream_t)stream, data,
                       g, axis, lpb, plan, sign);
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
