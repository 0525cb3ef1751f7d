/*
 * C-ABI of libhqc_amd.so: hand-written gfx950 kernels for the periodic GPW
 * (Gaussian + plane-wave) SCF path of hqc (hqc/pbc/solver.py, scf.py, gto.py,
 * hqc/tools/smearing.py, excor.py, hqc/pbc/potential.py).
 *
 * The reference is pure Python/JAX and exposes no FFI: its plugin surface is
 * the factory API make_lcao / make_pes / make_solver (hqc/pbc/lcao.py:11,
 * hqc/pbc/pes.py:10, hqc/pbc/solver.py:14).  These entry points are what a
 * host binding of that API calls, one per device operation of the reference's
 * jitted program; each comment names the reference operation it replaces.
 *
 * Conventions (all entry points):
 *  - every pointer is a DEVICE pointer owned by the caller; nothing is
 *    allocated or freed by the library unless stated;
 *  - `stream` is a hipStream_t (passed as void*); work is enqueued, never
 *    synchronised;
 *  - return 0 (HQC_OK) or an HQC_E_* code; no exception or exit crosses the ABI;
 *  - arrays carry a leading batch dimension B (independent geometries of
 *    identical shape), stored contiguously one geometry after another;
 *  - grids are g x g x g points, point index r = (ix*g + iy)*g + iz, the point
 *    coordinate is (ix, iy, iz) * L/g; "ld" is the padded row length of grid
 *    arrays (a multiple of 64, >= g^3; padding must hold zeros);
 *  - AO matrices are nao_pad x nao_pad (nao_pad = n_ao rounded up to 16, the
 *    padding rows/columns hold zeros), row-major, symmetric where stated;
 *  - complex arrays are interleaved (re, im) doubles.
 */
#ifndef HQC_AMD_H
#define HQC_AMD_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HQC_OK 0
#define HQC_E_ARG 1     /* bad shape / argument */
#define HQC_E_HIP 2     /* HIP launch or runtime error */
#define HQC_E_SIZE 3    /* size beyond what the kernel supports */

/* Library version (major*100 + minor). */
int hqc_version(void);

/* Row swaps over many tensors in one launch (no counterpart in the reference,
 * whose SCF keeps its whole batch): rows a[k] and b[k] (device int64, k < npairs
 * <= 65535, pairs disjoint) of each of ntensors tensors trade places; tensor t is
 * bases[t] (device) with row_bytes[t] bytes per leading-axis row; bases and
 * row_bytes are host arrays.  The batched SCF compacts the geometries still
 * iterating to a prefix of its per-geometry state with it.  HQC_E_SIZE beyond
 * HQC_SWAP_MAX_TENSORS tensors (the caller splits the list). */
#define HQC_SWAP_MAX_TENSORS 64
int hqc_swap_rows(void* const* bases, const long* row_bytes, int ntensors, const long* a,
                  const long* b, int npairs, void* stream);

/* Launch parameters of the calling host thread (no counterpart in the
 * reference, whose XLA programs are tuned by its compiler): set by the project's
 * autotuner (hqc_amd/tune.py) for a workload shape.  0 = built-in default.
 * They change how the work is split, not what is computed (partial sums of the
 * potential matrix are then added in another grouping).  Thread-local, so host
 * threads that drive different streams keep their own values.
 * hqc_set_tuning returns HQC_E_ARG for an unknown key or a value out of range;
 * hqc_get_tuning returns the value (-1 for an unknown key). */
#define HQC_TUNE_VMAT_BLOCKS 0       /* workgroups aimed at, FP64/FP32 V, n_ao <= 128 (2048) */
#define HQC_TUNE_VH_BLOCKS 1         /* the same for the split-FP16 V (2048) */
#define HQC_TUNE_VMAT_PARTS 2        /* grid chunks per geometry at most (64) */
#define HQC_TUNE_VMAT_TPW 3          /* upper-triangle tiles per wave, 1..8 (smallest that covers) */
#define HQC_TUNE_Q8_BLOCKS 4         /* int8-digit density: workgroups aimed at (2048) */
#define HQC_TUNE_Q8_MIN_TILES 5      /* int8-digit density: point tiles per wave at least (8) */
#define HQC_TUNE_FFT_PLANE_THREADS 6 /* 256, 512 or 1024 threads per y-z plane (by size) */
#define HQC_TUNE_VBIG_BLOCKS 7       /* workgroups aimed at, V for n_ao > 128 (1024) */
#define HQC_TUNE_FFT_TWO_BUFFER 8    /* 1: 3-D FFT passes between two LDS buffers instead of
                                        in place (the same outputs; kept for comparison) */
#define HQC_TUNE_FFT_IP_ITEMS 9      /* 1 or 2: work items per thread of the in-place FFT
                                        passes, when they fit (the fewest that fit) */
#define HQC_TUNE_EIGH_RECOMPUTE 10   /* 1: Jacobi recomputes pair positions and packed offsets
                                        per use instead of reading the round's table (the
                                        same results; for comparison) */
#define HQC_TUNE_XC_BLOCKS 11        /* workgroups per geometry of the LDA xc pass (8192 / B
                                        within 16..256; the partial sums follow it) */
#define HQC_TUNE_FORCE_POINTWISE 12  /* 1: the grid adjoint of the forces point by point
                                        (table reads per point) instead of along grid
                                        lines (the same sums in another order; for
                                        comparison) */
#define HQC_TUNE_AO_POINTWISE 13     /* 1: AO values point by point instead of along grid
                                        lines (the same values; for comparison) */
#define HQC_TUNE_COUNT 14
int hqc_set_tuning(int key, int value);
int hqc_get_tuning(int key);

/* 1-D periodic Gaussian-power tables.
 * Replaces eval_pbc_gaussian_power_x (hqc/pbc/gto.py:326-339) vmapped over
 * atoms and the 1-D grid (hqc/pbc/solver.py:211-214).
 *   P[b,A,d,j,a,l] = sum_{t=-tmax..tmax} (x - tL)^l exp(-alpha_a (x - tL)^2),
 *   x = j*L/g - xp[b,A,d], l = 0..nl-1.
 * xp (B,n,3) Bohr; alpha (na); P (B,n,3,g,na,nl). */
int hqc_gto_tables(const double* xp, int B, int n, int g, double L, int tmax,
                   const double* alpha, int na, int nl, double* P, void* stream);

/* AO values on the grid, the separable product of the 1-D tables.
 * Replaces eval_pbc_ao on Rmesh (hqc/pbc/gto.py:341-369, solver.py:217-219).
 *   phi[b, A*nloc+o, r] = sum_a coef[a,o] prod_d P[b,A,d,r_d,a,pow_d(o)]
 * comp[o] = -1 for s, 0/1/2 for p_x/p_y/p_z.  phi (B,nao_pad,ld) FP64 or FP32
 * (the _f32 variant rounds the FP64 value once). */
int hqc_ao_values_f64(const double* P, int B, int n, int g, int na, int nl,
                      const double* coef, const int* comp, int nloc, int nao_pad,
                      int ld, double* phi, void* stream);
int hqc_ao_values_f32(const double* P, int B, int n, int g, int na, int nl,
                      const double* coef, const int* comp, int nloc, int nao_pad,
                      int ld, float* phi, void* stream);
/* hqc_ao_values_f64 that also gives amax[b] = max |phi[b]| over the values it
 * writes (amax (B) zero-filled by the caller; the padding it leaves is zero in a
 * zero-filled phi).  The mixed set-up takes the FP16 pair scale from it before
 * its one pass over phi (hqc_phi_slice_i8_h2). */
int hqc_ao_values_f64_amax(const double* P, int B, int n, int g, int na, int nl,
                           const double* coef, const int* comp, int nloc, int nao_pad,
                           int ld, double* phi, double* amax, void* stream);

/* k-point tables and AO values (hqc/pbc/gto.py:371-420,
 * eval_pbc_gaussian_power_x_kpt / eval_pbc_ao_kpt):
 *   P[b,A,d,j,a,l] = sum_t (x - tL)^l exp(-alpha_a (x - tL)^2) exp(i k[b,d] t L)
 * complex128 (interleaved), kpt (B,3) in 1/Bohr.  hqc_ao_values_k writes the
 * complex AO values "stacked" into phi (B, 2*nao_pad, ld) FP64: rows
 * [0,nao_pad) Re phi, rows [nao_pad, 2 nao_pad) Im phi.  With this layout
 *   V = dV phi^H diag(v) phi: M = hqc_ao_vmat_f64 on the stacked rows,
 *     Re V = M_rr + M_ii, Im V = M_ri - M_ir;
 *   rho = Re sum_pq D_pq phi_p phi_q^*: hqc_ao_density_f64 with the real
 *     symmetric M = [[Re D, Im D], [-Im D, Re D]]. */
int hqc_gto_tables_k(const double* xp, const double* kpt, int B, int n, int g, double L,
                     int tmax, const double* alpha, int na, int nl, double* P, void* stream);
int hqc_ao_values_k(const double* P, int B, int n, int g, int na, int nl, const double* coef,
                    const int* comp, int nloc, int nao_pad, int ld, double* phi, void* stream);

/* AO values at arbitrary points, make_pbc_gto's eval_pbc_ao / eval_pbc_ao_kpt
 * (hqc/pbc/gto.py:341-369, 388-420) as used by make_slater
 * (hqc/pbc/slater.py:59-99): for each geometry b, point e and atom A, the 1-D
 * image sums of x_e - R_A over t = -tmax..tmax (tmax = rcut // L, times
 * e^{i k_d t L} when kpt != NULL).  xp (B,n,3), xe (B,ne,3) Bohr, kpt (B,3) or
 * NULL; out (B,ne,n*nloc) FP64, complex128 interleaved with kpt.  l_max <= 1
 * (nl <= 2). */
int hqc_ao_points(const double* xp, const double* xe, const double* kpt, int B, int n, int ne,
                  double L, int tmax, const double* alpha, int na, int nl, const double* coef,
                  const int* comp, int nloc, double* out, void* stream);

/* Geometry selection, common to the density and potential-matrix entry points:
 * bidx (B) int32 or NULL.  Entry b reads the AO values of geometry bidx[b] of
 * phi (which may hold more geometries than B, e.g. a whole SCF batch, so a
 * subset runs without copying phi); every other operand and every output is
 * indexed by b.  NULL means bidx[b] = b.
 *
 * Electron density on the grid and (optionally) X = D phi.
 * Replaces the 'pr,qr,pq->r' contraction of Exc_Vxc_integral
 * (hqc/pbc/solver.py:690) and the density used by hartree_rhoG (:627).
 *   rho[b,r] = sum_pq D[b,p,q] phi[b,p,r] phi[b,q,r]    (D symmetric)
 * FP64 FMAs on the vector ALUs with D as the scalar operand (X_p = sum_q
 * D_qp phi_q, then rho = sum_p phi_p X_p).  rho (B,ld); X (B,nao_pad,ld) or
 * NULL. */
int hqc_ao_density_f64(const double* phi, const int* bidx, const double* D, int B,
                       int nao_pad, int ld, double* rho, double* X, void* stream);
/* FP32 operands on v_mfma_f32_16x16x4f32 with FP32 accumulation over the AO
 * index; D is rounded to FP32; rho is written in FP64. */
int hqc_ao_density_f32(const float* phi, const int* bidx, const double* D, int B,
                       int nao_pad, int ld, double* rho, void* stream);

/* Density in orbital form, rho[b,r] = sum_i (sum_k At[b,k,i] phi[b,k,r])^2,
 * At = C sqrt(f) (nao_pad x m_pad, m_pad a multiple of 16, zero-padded) over
 * the orbitals with non-negligible occupation; equal to the D form above for
 * D = C f C^T.  Same reference operation (hqc/pbc/solver.py:690).
 * f64: FP64 FMAs on the vector ALUs with At read through scalar loads;
 * f32: FP32 phi and At on v_mfma_f32_16x16x4f32, rho accumulated in FP64. */
int hqc_orb_density_f64(const double* phi, const int* bidx, const double* At, int B,
                        int nao_pad, int m_pad, int ld, double* rho, void* stream);
int hqc_orb_density_f32(const float* phi, const int* bidx, const double* At, int B,
                        int nao_pad, int m_pad, int ld, double* rho, void* stream);
/* h2: from the FP16 pairs of hqc_phi_split_f16 (phi_scale as there); At is split
 * the same way with a per-geometry power-of-two scale, psi = At^T phi takes
 * hi*hi + hi*lo + lo*hi on v_mfma_f32_16x16x32_f16 with FP32 accumulation
 * (relative error ~2^-22 of sum_k |At phi|), rho accumulated in FP64. */
int hqc_orb_density_h2(const unsigned* ph, const int* bidx, const double* At, int B,
                       int nao_pad, int m_pad, int ld, double phi_scale, double* rho,
                       void* stream);
/* q8: from int8 digits with exact integer accumulation (Ozaki-style slicing).
 * hqc_phi_slice_i8 cuts phi (G,nao_pad,ld) FP64, once per geometry, into
 * nslice (3..5) base-256 int8 digits per value under a per-grid-point
 * power-of-two exponent: q (G, ld*nslice*nao_pad) bytes in tiles of 16 points
 * [tile][digit][AO block of 16][point][AO], pe (G,ld) the exponents.
 * hqc_orb_density_q8 cuts At per call into top + 1 digits per orbital
 * (per-orbital exponent), forms every digit product of level <= top (4 or 5;
 * 3 with nuse 3 or 4, 9 or 10 products, for the coarser SCF iterations)
 * on v_mfma_i32_16x16x64_i8 with int32 accumulation (exact), combines the
 * levels in FP64 and sums psi^2 in FP64.  Error: the fixed phi rounding,
 * 2^(1-6-8(nuse-1)) of max_k |phi_k(r)| per value (a perturbation of the basis
 * that does not change between calls), plus ~2^(-6-8 top) of max|At_i|
 * max|phi(r)| per AO from the coefficient digits (2^-46 at top 5; top 4 takes
 * 14 instead of 18 digit products with 4 phi digits).  nao_pad <= 128
 * (HQC_E_SIZE beyond).  nuse (3..nslice) of the nslice stored digits enter the
 * density (fewer: a coarser fixed basis and fewer bytes read).
 * amax (G) or NULL: if given (zero-filled by the caller), receives max |phi|
 * over each geometry's nao_pad x ld values, exactly, from the same pass. */
int hqc_phi_slice_i8(const double* phi, int G, int nao_pad, int ld, int nslice,
                     signed char* q, short* pe, double* amax, void* stream);
/* hqc_phi_slice_i8 (without amax) that in the same read of phi also writes h2
 * (G,nao_pad,ld): the FP16 pairs of scale*phi, the values hqc_phi_split_f16 gives
 * (scale = 2^k, |scale phi| <= 2^14). */
int hqc_phi_slice_i8_h2(const double* phi, int G, int nao_pad, int ld, int nslice,
                        double scale, signed char* q, short* pe, unsigned* h2, void* stream);
int hqc_orb_density_q8(const signed char* q, const short* pe, const int* bidx,
                       const double* At, int B, int nao_pad, int m_pad, int ld, int nslice,
                       int nuse, int top, double* rho, void* stream);

/* Grid potential matrix V[b] = dV * phi diag(v[b]) phi^T (symmetric, full).
 * Replaces the 'pr,qr,r->pq' contraction (hqc/pbc/solver.py:693) and, through
 * the FFT-Poisson equivalence, the rhoG contractions for V_ep (:427,470) and
 * J (:628).  FP64 MFMA; per-chunk partial tiles are reduced in a fixed order
 * (deterministic).  work: device scratch of hqc_ao_vmat_workspace bytes. */
int hqc_ao_vmat_workspace(int B, int nao_pad, int ld, int prec, size_t* bytes);
int hqc_ao_vmat_f64(const double* phi, const int* bidx, const double* v, int B, int nao_pad,
                    int ld, double dV, double* V, void* work, size_t work_bytes, void* stream);
/* FP32 phi and v on v_mfma_f32_16x16x4f32 with FP32 accumulation inside a
 * block, FP64 accumulation across blocks; V written in FP64. */
int hqc_ao_vmat_f32(const float* phi, const int* bidx, const double* v, int B, int nao_pad,
                    int ld, double dV, double* V, void* work, size_t work_bytes, void* stream);
/* Allocation-free form of hqc_ao_vmat_f64 (no counterpart in the reference):
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops).  Same arguments and result without the
 * workspace; the partial tiles live in a device array of the library of
 * hqc_ao_vmat_scratch_bytes() (32 MiB) per device, held until the process
 * ends.  Results equal hqc_ao_vmat_f64's bit for bit while
 * B * n_tiles * n_chunks * 2 KiB fits in it (n_tiles = (nao_pad/16)(nao_pad/16+1)/2,
 * n_chunks as hqc_ao_vmat_workspace chooses); beyond that fewer grid chunks per
 * geometry are taken (the same sum in another order); HQC_E_SIZE when even one
 * chunk per geometry does not fit (B * n_tiles * 2 KiB > 32 MiB).  Calls that
 * overlap on one device (several streams or host threads) share the array and
 * race; calls in sequence on one stream do not. */
size_t hqc_ao_vmat_scratch_bytes(void);
int hqc_ao_vmat_f64_noalloc(const double* phi, const int* bidx, const double* v, int B,
                            int nao_pad, int ld, double dV, double* V, void* stream);
/* Split-precision form of the same contraction (nao_pad <= 128).
 * hqc_phi_split_f16: out[i] = (hi, lo) FP16 pair (hi in the low half) with
 *   scale*phi[i] = hi + lo to 22 significant bits; scale = 2^k, |scale phi| <= 2^14.
 * hqc_ao_vmat_h2: V[b] = dV phi diag(v[b]) phi^T from those pairs; v phi is split
 *   the same way in LDS with a per-geometry power-of-two scale from vmax[b] =
 *   max_r |v[b,r]| (device array, or NULL: computed here), and each tile takes
 *   hi*hi + hi*lo + lo*hi on v_mfma_f32_16x16x32_f16 with FP32 accumulation
 *   (relative error ~2^-22 of sum |phi phi v|).  work: hqc_ao_vmat_h2_workspace bytes. */
int hqc_phi_split_f16(const double* phi, long n, double scale, unsigned* out, void* stream);
int hqc_ao_vmat_h2_workspace(int B, int nao_pad, int ld, size_t* bytes);
int hqc_ao_vmat_h2(const unsigned* ph, const int* bidx, const double* v, const double* vmax,
                   int B, int nao_pad, int ld, double phi_scale, double dV, double* V,
                   void* work, size_t work_bytes, void* stream);
/* Operands of the mixed-precision V step (no counterpart in the reference,
 * whose V is FP64 throughout): for each geometry b, vmax[b] = max_r |v[b,r]|,
 * dvmax[b] = max_r |v[b,r] - va[b,r]|, and vv[b,:] = v[b,:] where direct[b]
 * (one byte per geometry, nonzero = true), v[b,:] - va[b,:] elsewhere.  One pass
 * over v and va; values as from the separate FP64 operations (NaN propagates
 * into the maxima).  v, va, vv (B,ld), 16-byte aligned rows, ld even. */
int hqc_mixed_v_prep(const double* v, const double* va, const unsigned char* direct, int B,
                     int ld, double* vv, double* vmax, double* dvmax, void* stream);

/* Batched in-place 3-D complex FFT of B grids g^3 (any g <= 128; mixed-radix
 * Stockham passes in LDS, one launch per axis).  sign = -1 forward,
 * +1 inverse, unnormalised.  Replaces jnp.fft.fft (hqc/pbc/solver.py:412)
 * in the FFT-Poisson restatement. */
int hqc_fft3d_c2c_f64(double* data, int B, int g, int sign, void* stream);
/* The same transform of complex64 grids (interleaved floats; twiddles formed in
 * FP64 and rounded once): the FP32 pair-density transforms of the exchange
 * (hqc/pbc/solver.py:412, 430-431 in the reference, FP64 there). */
int hqc_fft3d_c2c_f32(float* data, int B, int g, int sign, void* stream);

/* Hartree-Fock exchange pair densities, two orbitals per complex grid:
 *   z[m, p, r] = phi[b, p, r] * (psi[b, i, r] + i psi[b, j, r]),
 * table[m] = (b, i, j) with j = -1 for a lone orbital; psi rows carry the
 * weight sqrt(f_i) dV.  z is (mp, nao, g^3) complex128, ready for
 * hqc_fft3d_c2c_f64(sign -1).  Replaces the AO-pair densities rhoG that feed
 * eris (hqc/pbc/solver.py:412-431) on the exchange route K = sum_rs D_rs eris. */
int hqc_pair_pack(const double* phi, const double* psi, const int* table, int mp, int nao,
                  int nao_pad, int nmo, int g, int ld, double* z, void* stream);
/* _c64: z as complex64 (each FP64 product rounded once), for hqc_fft3d_c2c_f32. */
int hqc_pair_pack_c64(const double* phi, const double* psi, const int* table, int mp, int nao,
                      int nao_pad, int nmo, int g, int ld, float* z, void* stream);
/* hqc_pair_pack_fft: hqc_pair_pack followed by the 3-D transform of sign `sign`
 * (hqc_fft3d_c2c_f64) in one pass less: the pair densities are formed while the
 * transform's y-z planes are loaded, so z is written once, by the transform (the
 * same products, the same bits).  Grid (m, p) lands at z + (m z_rows + p) z_stride
 * (complex elements; z_rows >= nao, z_stride >= g^3; other elements untouched):
 * z_rows = nao, z_stride = g^3 is hqc_pair_pack's layout, and z_rows = nao_pad with
 * zeroed padding makes z, read as real rows [Re Z | Im Z ...] of 2 z_stride values,
 * an operand of hqc_ao_vmat_* (then K_pq = sum_G VG Re[Z_p(G) conj Z_q(G)] over all
 * G, which equals the two orbitals' sum_G VG (Re[A_a A_a*] + Re[A_b A_b*]): the
 * terms odd in G cancel over the symmetric G-set, so no split is needed).
 * HQC_E_SIZE where the y-z plane does not fit the LDS (g > 98) or mp * nao > 65535:
 * use hqc_pair_pack and the transform there.  _c64: the same with hqc_pair_pack_c64
 * and hqc_fft3d_c2c_f32.  (The replacement of hqc_pair_pack, hqc/pbc/solver.py:412-431.) */
int hqc_pair_pack_fft(const double* phi, const double* psi, const int* table, int mp, int nao,
                      int nao_pad, int nmo, int g, int ld, double* z, int z_rows, int z_stride,
                      int sign, void* stream);
int hqc_pair_pack_fft_c64(const double* phi, const double* psi, const int* table, int mp,
                          int nao, int nao_pad, int nmo, int g, int ld, float* z, int z_rows,
                          int z_stride, int sign, void* stream);

/* Split each transformed pair grid Z = FFT(a + i b) into the real row blocks
 * M[2m] = [Re A_a(H) | Im A_a(H)] and M[2m+1] = [Re A_b(H) | Im A_b(H)] over the
 * half G-set H (centred order, columns < (g^3-1)/2); M is (2 mp, nao_pad, ld),
 * rows >= nao and columns >= g^3-1 written as zero.  Then
 * K = sum_m M_m diag(w) M_m^T with w = 2 VG on both halves (hqc_ao_vmat_*,
 * dV = 1) is the G != 0 part of the exchange matrix
 * (hqc/pbc/solver.py:430-431, 604-614). */
int hqc_pair_split_f64(const double* z, int mp, int nao, int nao_pad, int g, int ld, double* M,
                       void* stream);
int hqc_pair_split_f32(const double* z, int mp, int nao, int nao_pad, int g, int ld, float* M,
                       void* stream);
/* _c64: from complex64 transforms (hqc_pair_pack_c64 + hqc_fft3d_c2c_f32) to FP32 M. */
int hqc_pair_split_c64(const float* z, int mp, int nao, int nao_pad, int g, int ld, float* M,
                       void* stream);

/* Local (electron-proton) potential in G space, FFT order:
 *   out[b,G] = -VG[G] * sum_A exp(-i G.R_A)   (hqc/pbc/solver.py:406-407).
 * xp (B,n,3) Bohr; VG (g^3) FFT order; out complex (B,g^3). */
int hqc_vloc_G(const double* xp, int B, int n, int g, double L, const double* VG,
               double* out, void* stream);

/* Real grid (B,ld) -> complex (B,g^3) scaled by s: out = (s*in, 0). */
int hqc_real_to_complex(const double* in, int B, int g, int ld, double s, double* out,
                        void* stream);
/* Two real grids per complex grid: out ((B+1)/2, g^3) = s (in[2p] + i in[2p+1]),
 * imaginary part 0 for the last grid of an odd B.  With a real, even G-space
 * kernel (VG) the FFT, multiply and inverse FFT keep the two results in Re and
 * Im: one transform pair per two geometries. */
int hqc_real_pair_to_complex(const double* in, int B, int g, int ld, double s, double* out,
                             void* stream);
/* Complex (B,g^3) -> real part into (B,ld). */
int hqc_complex_real(const double* in, int B, int g, int ld, double* out, void* stream);
/* Poisson kernel in G space: c[b,G] *= VG[G] (VG (g^3) shared, or (B,g^3) one
 * row per geometry when vg_batched != 0, for batches of different cell sizes);
 * if saved != NULL the incoming c is copied there first (kept for forces). */
int hqc_poisson_mul(double* c, const double* VG, int B, int g, int vg_batched, double* saved,
                    void* stream);

/* Pointwise exchange-correlation and energy densities.
 * Replaces exc_functional / vxc_functional (hqc/pbc/solver.py:669-678,
 * hqc/tools/excor.py) with closed-form derivatives.
 *   xc_id: 0 none (HF), 1 lda+vwn, 2 lda only, 3 lda+hl71, 4 lda+bh72, 5 lda+gl77
 * rho (B,ld); vh complex (B,g^3) from the inverse FFT (real part = v_H), or
 * with vh_paired the ((B+1)/2, g^3) grids of hqc_real_pair_to_complex (v_H of
 * geometry b = Re (b even) or Im (b odd) of grid b/2);
 * vout (B,ld) = v_H + v_xc (padding left untouched);
 * part (B, nblk, 2) per-block sums of rho*eps_xc and rho*v_H; nblk is
 * returned by hqc_xc_blocks for the same B on the same host thread (it follows
 * the batch size: 8192 workgroups in all, 16..256 per geometry). */
int hqc_xc_blocks(int B, int g);
int hqc_xc_eval(const double* rho, const double* vh, int vh_paired, int B, int g, int ld,
                int xc_id, double* vout, double* part, void* stream);

/* GGA (PBE) on the grid; not in the reference (which accepts only LDA
 * exchange, hqc/tools/excor.py:14-17).  Spectral derivatives on the odd grid:
 *   hqc_grad_G: out (B,3,g^3) complex = scale * i G_d rhoG (inverse FFT ->
 *     grad rho with scale = 1/Omega when rhoG = dV FFT(rho));
 *   hqc_xc_gga_eval: rho (B,ld), grad (B,3,g^3) complex (real part used), vh as
 *     for hqc_xc_eval; vout (B,ld) = v_H + de/drho, w (B,3,g^3) complex =
 *     2 de/dsigma grad rho, part (B, hqc_gga_blocks, 2) sums of e and rho v_H;
 *   hqc_div_G: out (B,g^3) = scale * sum_d i G_d wG[b,d] (forward FFT of w in,
 *     inverse FFT of out with scale 1/g^3 -> div w); v_xc = de/drho - div w;
 *   hqc_pbe_points: out (n,3) = (e, de/drho, de/dsigma) at given points.
 * Lb (B) gives each geometry its own cell length (NULL: L for all). */
int hqc_grad_G(const double* rhoG, int B, int g, double L, const double* Lb, double scale,
               double* out, void* stream);
int hqc_div_G(const double* wG, int B, int g, double L, const double* Lb, double scale,
              double* out, void* stream);
int hqc_gga_blocks(int g);
int hqc_xc_gga_eval(const double* rho, const double* grad, const double* vh, int B, int g,
                    int ld, double* vout, double* w, double* part, void* stream);
int hqc_pbe_points(const double* rho, const double* sigma, int n, double* out, void* stream);

/* Chemical potential and occupations (hqc/tools/smearing.py:4-173), one
 * wavefront per geometry, reproducing the reference loop semantics including
 * the bisection return rule of smearing.py:102.
 * w (B,nmo) Hartree ascending; method 0 fermi / 1 gauss; search 0 bisect /
 * 1 newton; occ (B,nmo); mu (B). */
int hqc_occupations(const double* w, int B, int nmo, int nelec, double sigma, int method,
                    int search, int max_cycle, double tol, double mu_shift, double* occ,
                    double* mu, void* stream);

/* Batched symmetric eigensolver, cyclic two-sided Jacobi in LDS, n <= 110.
 * Replaces jnp.linalg.eigh of the orthogonalised Fock matrix
 * (hqc/pbc/scf.py:69,138,198, hqc/pbc/solver.py:741).
 * A (B,n,n) symmetric; pairs ((npad-1) x npad ints, npad = n rounded up to
 * even) is a round-robin pairing schedule: every row lists npad/2 disjoint
 * pairs and every pair occurs once per sweep; NULL selects the circle method
 * computed on the device (index 0 fixed, the rest rotated one place per round).
 * Rotations with |a_pq| <= tol ||A||_F / n are skipped (threshold Jacobi).
 * Stops when the off-diagonal Frobenius norm <= tol * ||A||_F or after
 * max_sweeps.  w (B,n) ascending;
 * V (B,n,n) eigenvectors as columns; sweeps (B) or NULL.
 * hqc_eigh_lds_bytes(n) is the LDS one block needs. */
size_t hqc_eigh_lds_bytes(int n);
int hqc_eigh_jacobi(const double* A, int B, int n, const int* pairs, int max_sweeps, double tol,
                    double* w, double* V, int* sweeps, void* stream);

/* Analytic periodic overlap and kinetic matrices by lattice sums
 * (hqc/pbc/solver.py:233-389, the s/p derivative trick of :307-341).
 * xp (B,n,3) Bohr; lattice (nT,3); alpha (na); coef (na,nloc); comp (nloc).
 * S,T (B,n_ao,n_ao).  If dS/dT are non-NULL also the first-argument
 * derivatives dS[b,p,q,c] = dS_pq/dR_{atom(p),c} (B,n_ao,n_ao,3). */
int hqc_overlap_kinetic(const double* xp, int B, int n, const double* lattice, int nT,
                        const double* alpha, int na, const double* coef, const int* comp,
                        int nloc, double* S, double* T, double* dS, double* dT,
                        void* stream);

/* The same for bases of s functions only (comp all -1, e.g. gth-szv, gth-dzv) with
 * na = 4 primitives and nloc <= 2 functions per atom (HQC_E_SIZE otherwise): one
 * thread per atom pair forms each lattice exponential once for all its contracted
 * pairs; every entry is summed in hqc_overlap_kinetic's order (equal to rounding). */
int hqc_overlap_kinetic_s(const double* xp, int B, int n, const double* lattice, int nT,
                          const double* alpha, int na, const double* coef, int nloc, double* S,
                          double* T, double* dS, double* dT, void* stream);

/* k-point overlap and kinetic matrices, eval_overlap_kinetic_kpt
 * (hqc/pbc/solver.py:257-305): every lattice term carries exp(-i k.T)
 * (:270-271).  kpt (B,3) 1/Bohr; S,T (B,n_ao,n_ao) complex128 interleaved;
 * dS, dT (B,n_ao,n_ao,3) complex128 or NULL (first-argument derivatives). */
int hqc_overlap_kinetic_k(const double* xp, const double* kpt, int B, int n,
                          const double* lattice, int nT, const double* alpha, int na,
                          const double* coef, const int* comp, int nloc, double* S, double* T,
                          double* dS, double* dT, void* stream);

/* Ewald proton-proton sum (hqc/pbc/potential.py:35-67,103-130).
 * xfrac (B,npair,3) minimum-image fractional separations; Gv (nG,3) integer
 * vectors with weights gk (nG); out (B,npair,4) = (psi, dpsi/dr_x,y,z). */
int hqc_ewald_pairs(const double* rij, int B, int npair, const double* Gv, const double* gk,
                    int nG, double kappa, double g0, double* out, void* stream);
/* The reciprocal sums of hqc_ewald_pairs summed over the pairs, by structure
 * factors (n N_G phases per geometry instead of n(n-1)/2 N_G): frac (B,n,3)
 * fractional coordinates, E (B) = sum_{i<j} sum_G gk cos(2 pi G.r_ij),
 * grad (B,n,3) = d E / d frac_i; work holds B*nG*2 doubles.  Same reference
 * sum (hqc/pbc/potential.py:35-67) as the pair kernel, in another order. */
int hqc_ewald_recip(const double* frac, int B, int n, const double* Gv, const double* gk, int nG,
                    double* E, double* grad, double* work, void* stream);

/* Forces: grid adjoint of the AO contractions (new; the reference has none).
 *   F[b,A,c] = dV * sum_r sum_{p in A} Z_p(r) d phi_p(r)/d r_c,
 *   Z = 2 v(r) X_p(r) + Zx_p(r)  (Zx may be NULL)
 * grad phi is recomputed from tables P built with nl = l_max + 2 powers and at
 * least 3 (nl = 3 for s and p bases; HQC_E_ARG below 3); alpha (na) are the
 * primitive exponents.
 * part (B, n, nblk, 3), nblk = hqc_force_blocks(g): partial sums over disjoint
 * parts of the work, F = their sum over the block axis (for g <= 128: block
 * o*C + k, C = nblk / nloc, holds AO o's sum over the k-th of C ranges of grid
 * lines, zero past nloc*C; HQC_TUNE_FORCE_POINTWISE or g > 128: blocks of grid
 * points).  HQC_E_SIZE for nloc > nblk. */
int hqc_force_blocks(int g);
int hqc_grid_forces(const double* P, int B, int n, int g, int na, int nl, const double* alpha,
                    const double* coef, const int* comp, int nloc, int nao_pad, int ld,
                    const double* X, const double* v, const double* Zx, double dV,
                    double* part, void* stream);

/* Grid adjoint at a k-point (new): complex tables P (B,n,3,g,na,nl)
 * complex128 from hqc_gto_tables_k with nl = l_max + 2; X, Zx stacked
 * (B, 2*nao_pad, ld) as from hqc_ao_density_f64 on stacked rows;
 *   F[b,A,c] = dV sum_r sum_{p in A} (Z_re d Re phi_p/dr_c + Z_im d Im phi_p/dr_c),
 *   Z = 2 v X + Zx. */
int hqc_grid_forces_k(const double* P, int B, int n, int g, int na, int nl, const double* alpha,
                      const double* coef, const int* comp, int nloc, int nao_pad, int ld,
                      const double* X, const double* v, const double* Zx, double dV,
                      double* part, void* stream);

/* Structure-factor derivative of the local potential energy:
 *   dE[b,A,c] = Re sum_G i G_c VG(G) exp(-i G.R_A) conj(rho~[b,G]). */
int hqc_vloc_force(const double* xp, int B, int n, int g, double L, const double* VG,
                   const double* rhoG, double* dE, void* stream);

/* SCF driver: the batched Gamma-point restricted Kohn-Sham SCF of
 * make_solver(..., dft=True) (hqc/pbc/solver.py:705-767, the DIIS loop of
 * hqc/pbc/scf.py:93-242), run from C on the entry points above and rocBLAS, for
 * hosts without Python (the reference has no FFI: its plugin surface is the
 * Python factory, which hqc_amd.pbc.make_solver mirrors).  Scope: FP64, LDA
 * functionals (xc_id 1..5 of hqc_xc_eval), n * nloc <= 110, DIIS on, one cell
 * length per context.  The cell's arrays are host memory, copied at create.
 *   g: grid points per axis, the reference's odd grid (hqc_scf_grid);
 *   alpha (na), coef (na, nloc), comp (nloc): the normalised basis with the
 *     Y_lm factors folded in, as hqc_ao_values takes them;
 *   rcut: lattice-sum cut-off of S and T in Bohr (24 in the reference);
 *   smearing 0: aufbau occupations (smearing.py:171); 1: smearing_method
 *     0 fermi / 1 gauss at sigma (Hartree), search 0 bisect / 1 newton;
 *   tol, max_cycle, diis_space, diis_start_cycle, diis_damp as make_solver's.
 * hqc_scf_run: xp (B, n, 3) Bohr on the device (the reference's xp * rs);
 * device outputs, any of them NULL to skip: mo (B, n_ao, n_ao), dm (B, n_ao,
 * n_ao), bands (B, n_ao) Ry, E, Ki, Vep, Vee (B) Ry, Se (B), converged and
 * n_iter (B) int32 -- make_solver's 9-tuple plus the iteration counts.  The
 * call allocates its work memory and synchronises the stream before it returns.
 * hqc_scf_grid: round(L / grid_length / 2) * 2 + 1 with ties to even
 * (solver.py:159). */
typedef struct hqc_scf_cell {
  int n;
  double L;
  int g;
  int na, nloc, nl;
  const double* alpha;
  const double* coef;
  const int* comp;
  double rcut;
  int xc_id;
  int smearing, smearing_method;
  double sigma;
  int search, search_cycle;
  double search_tol;
  double tol;
  int max_cycle;
  int diis_space, diis_start_cycle;
  double diis_damp;
} hqc_scf_cell;
typedef struct hqc_scf_ctx hqc_scf_ctx;
int hqc_scf_grid(double L, double grid_length);
int hqc_scf_create(const hqc_scf_cell* cell, hqc_scf_ctx** ctx);
int hqc_scf_run(hqc_scf_ctx* ctx, const double* xp, int B, double* mo, double* dm,
                double* bands, double* E, double* Ki, double* Vep, double* Vee, double* Se,
                int* converged, int* n_iter, void* stream);
/* hqc_scf_forces: the same SCF, then the energy and forces of make_pes(...,
 * dft=True).energy_and_forces (hqc/pbc/pes.py:10-154 with the forces this package
 * adds): E_tot (B) Ry = E + V_pp (the Ewald sum of potential.py:35-67 with kappa,
 * Gmax), F (B, n, 3) Ha/Bohr = -dA/dR, A = E - sigma S + V_pp (Pulay, grid
 * adjoint, local potential and Ewald terms); converged, n_iter as above.  F is
 * required; Gaussian smearing is refused (HQC_E_ARG: its forces are not the
 * gradient of A). */
int hqc_scf_forces(hqc_scf_ctx* ctx, const double* xp, int B, double kappa, int Gmax,
                   double* E_tot, double* F, int* converged, int* n_iter, void* stream);
void hqc_scf_destroy(hqc_scf_ctx* ctx);

#ifdef __cplusplus
}
#endif
#endif /* HQC_AMD_H */
