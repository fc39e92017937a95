/*
 * fastsht.h -- C ABI of libfastsht (MI355X / gfx950).
 *
 * Batched HEALPix-RING spherical harmonic transforms in FP64: many maps at
 * once, every transform stage on the GPU.  Drop-in boundary for the
 * reference's Fortran module `fastSHT` (liuhao-cn/fastSHT, wrapped by f90wrap
 * and driven from SHT.py); each entry point names the reference interface it
 * replaces.
 *
 * Layouts (all Fortran order, float64), as in the reference:
 *   maps  (npix, nsim)             map k at maps + k*npix, npix = 12*nside^2, RING order
 *   alms  (nsim, lmax+1, lmax+1)   Re a_lm at [k, l, m] (l >= m),
 *                                  Im a_lm at [k, l-m, lmax+1-m] (m >= 1)
 *
 * Pointers are host memory (loc = FSHT_HOST; the call copies through HBM and
 * returns when results are in host memory) or device memory (loc =
 * FSHT_DEVICE; the call is asynchronous on the handle's stream, see
 * fsht_get_stream / fsht_synchronize).
 *
 * Every function returns FSHT_OK (0) or a negative status; fsht_last_error()
 * describes the failure.  A handle must not be used from two host threads at
 * once; distinct handles are independent.
 */
#ifndef FASTSHT_H
#define FASTSHT_H

#ifdef __cplusplus
extern "C" {
#endif

#define FSHT_OK 0
#define FSHT_ERR_ARG (-1)   /* invalid argument */
#define FSHT_ERR_NOMEM (-2) /* device allocation failed */
#define FSHT_ERR_HIP (-3)   /* HIP runtime error */
#define FSHT_ERR_STATE (-4) /* call not valid for this handle's configuration/state */

#define FSHT_HOST 0
#define FSHT_DEVICE 1

typedef struct fsht_handle fsht_handle;

/* Library identification string. */
const char* fsht_version(void);

/* Message of the last failure on this handle (or of the calling thread when h is NULL). */
const char* fsht_last_error(const fsht_handle* h);

/* Create a transform configuration on the current HIP device: ring geometry,
 * ring-skip table, Legendre start seeds and workspace stay resident.
 * Replaces SHT.__init__ (SHT.py:17-62) -> sht_data_alloc + sht_set_data
 * (src/sht_data_init.f90:2-311).  nside: power of two <= 2048;
 * 0 <= lmax <= 4*nside; pol != 0 enables the spin-2 calls. */
int fsht_create(int nside, int lmax, int nsim, int niter, int pol, fsht_handle** out);

/* Release a handle.  Replaces sht_data_free (src/sht_data_init.f90:319-372). */
int fsht_destroy(fsht_handle* h);

/* The handle's HIP stream (hipStream_t) and a blocking wait on it. */
int fsht_get_stream(fsht_handle* h, void** stream);
/* Enqueue this handle's work on a caller stream (hipStream_t; NULL is the null
 * stream, as torch's default stream is).  A handle starts on a stream of its
 * own, which fsht_get_stream returns before the first switch; handles on
 * different non-null streams are not ordered with each other.  The caller
 * keeps ownership of the stream. */
int fsht_set_stream(fsht_handle* h, void* stream);
int fsht_synchronize(fsht_handle* h);

/* Number of maps of the following calls: 1 <= nsim <= the nsim the handle was
 * created with (its capacity).  Arrays passed to later calls hold nsim maps
 * (leading dimension nsim); work on the capacity's unused 16-map column tiles
 * is skipped.  No counterpart in the reference (one fixed nsim per
 * configuration): it lets a serving loop pack requests of one shape into one
 * resident handle.  Drops alms kept on the device by fsht_t2alm(alms = NULL). */
int fsht_set_batch(fsht_handle* h, int nsim);

/* out[0..n) = {nside, lmax, nsim, niter, pol, npix, nsim padded to 16, m-major
 * buffer size, precision, device bytes held for mixed precision, and the
 * mixed-precision slice counts: k_lo, k_hi, damp_min, k_out, damp_min_a}.  In
 * mixed precision the first analysis of a t2alm(niter) and a residual
 * synthesis whose error the later Jacobi iterations damp at least damp_min
 * times (analysis k: niter - k, synthesis k: niter - k - 1), and a later
 * analysis damped at least damp_min_a times, use k_lo slices per operand, the
 * others k_hi; alm2t and t2alm with niter = 0 use k_out; then the capacity
 * (the nsim the handle was created with); then, for spin 0 (out[16..26)) and
 * spin 2 (out[26..36)), the slices of every pass as a table: out[b + 3 k + d]
 * for pass kind k (0: the first analysis, 1: a residual synthesis, 2: a
 * residual analysis) whose error d = min(2, damping) later iterations damp,
 * and out[b + 9] for alm2t / eb2qu and an unrefined analysis; then the device
 * bytes the mixed-precision state takes once built, planned on the host
 * whether or not it is built yet: out[36] the sliced operands (set up by
 * fsht_set_precision), out[37] the temperature tables, out[38] the spin-2
 * exponents and the two range buffers (lets a caller budget a handle's
 * memory before switching it to mixed precision). */
int fsht_info(const fsht_handle* h, long long* out, int n);

#define FSHT_PREC_FP64 0  /* every Legendre contraction in FP64 (the default) */
#define FSHT_PREC_MIXED 1 /* Legendre contractions on int8-slice MFMA, refined in FP64 */

/* Precision of the Legendre contractions of this handle's transforms
 * (fsht_t2alm, fsht_t2alm_iter, fsht_alm2t, fsht_qu2eb, fsht_eb2qu; fsht_qu2e
 * and fsht_qu2b stay FP64).  FSHT_PREC_MIXED evaluates each contraction as
 * exact int8 products of 7-bit fixed-point digits (4 to 6 per operand,
 * v_mfma_i32_32x32x32_i8) and keeps the ring FFTs, the Jacobi residual and
 * the alm accumulation in FP64; t2alm(niter) and qu2eb(niter) then match the
 * FP64 result to about 1e-10 (max |diff| / max |alm|), alm2t and eb2qu to
 * about 1e-11.  No counterpart in the reference, whose GEMMs
 * (src/sht_core.f90:30-168) are FP64 only.  Switching a temperature-only
 * handle to FSHT_PREC_MIXED builds its sliced Legendre tables (about 12 bytes
 * per (m, l, ring) term plus sliced-operand workspace) before returning; a
 * spin-2 capable handle builds its temperature tables at its first mixed
 * temperature transform, and at its first mixed spin-2 transform the F1/F2
 * exponents and two buffers that each pass generates the F1/F2 tables of a
 * range of orders into (each at most 6 GiB and 1/16 of the handle's FP64
 * workspace, at least 256 MiB or one order's tables). */
int fsht_set_precision(fsht_handle* h, int precision);

/* The ring-skip table used (first kept north ring per m, m = 0..lmax), as read
 * from SHT_ring_skip_table/NNNN.bin at SHT.py:31-32. */
int fsht_ring_tab(const fsht_handle* h, int* out);

/* Temperature analysis maps -> alms, with niter Jacobi refinements.
 * niter = 0 replaces t2alm (src/sht_main.f90:38-89); niter > 0 replaces
 * t2alm_iter_old (src/sht_main.f90:276-367), called from SHT.t2alm
 * (SHT.py:72-107).  alms = NULL keeps the result on the device for
 * fsht_get_alm (the reference's get_alm=False path, SHT.py:88-95). */
int fsht_t2alm(fsht_handle* h, const double* maps, double* alms, int niter, int loc);

/* Temperature analysis with the "immediate update" refinement: replaces
 * t2alm_iter (src/sht_main.f90:161-264), selected in SHT.t2alm by
 * method != 'old' (SHT.py:100-101).  Each of the niter sweeps re-analyses,
 * for every order m, the residual of the maps with only the synthesised
 * aliases m' <= m of its ring frequency subtracted (the reference's order-m
 * scatter, m ascending); niter = 0 equals fsht_t2alm.  alms = NULL keeps the
 * result on the device as in fsht_t2alm. */
int fsht_t2alm_iter(fsht_handle* h, const double* maps, double* alms, int niter, int loc);

/* Fetch the alms kept by fsht_t2alm(..., alms=NULL).  Replaces get_alm
 * (src/sht_main.f90:1294-1302) / SHT.get_alms (SHT.py:214-227). */
int fsht_get_alm(fsht_handle* h, double* alms, int loc);

/* Temperature synthesis alms -> maps.  Replaces alm2t (src/sht_main.f90:96-155)
 * / SHT.alm2t (SHT.py:109-122). */
int fsht_alm2t(fsht_handle* h, const double* alms, double* maps, int loc);

/* Spin-2 analysis Q,U -> E,B with niter Jacobi refinements.  niter = 0
 * replaces QU2EB (src/sht_main.f90:383-436); niter > 0 replaces
 * qu2eb_iter_old (src/sht_main.f90:563-673); SHT.qu2eb (SHT.py:124-144). */
int fsht_qu2eb(fsht_handle* h, const double* Q, const double* U, double* E, double* B, int niter, int loc);

/* E only / B only.  Replace QU2E (src/sht_main.f90:955-1005) and QU2B
 * (src/sht_main.f90:1016-1064); SHT.qu2e / SHT.qu2b (SHT.py:146-176). */
int fsht_qu2e(fsht_handle* h, const double* Q, const double* U, double* E, int loc);
int fsht_qu2b(fsht_handle* h, const double* Q, const double* U, double* B, int loc);

/* Spin-2 synthesis E,B -> Q,U.  Replaces EB2QU (src/sht_main.f90:1079-1147)
 * / SHT.eb2qu (SHT.py:179-190). */
int fsht_eb2qu(fsht_handle* h, const double* E, const double* B, double* Q, double* U, int loc);

/* EB-leakage cleaning of polarised maps on a masked sky.  Replaces fix_eb
 * (src/sht_main.f90:801-941) behind SHT.fix_eb (SHT.py:192-202).
 * Q and U are masked in place (Q *= mask, U *= mask), as the reference does;
 * mask has npix entries.  Steps: (E, B) = qu2eb(masked Q, U, niter); the
 * E-only synthesis of E, masked again, analysed again gives the leakage
 * template B'; Bmap = alm2t(B) - a0 - a1 * alm2t(B') with (a0, a1) the
 * least-squares fit per map over the pixels where mask == 1 (Bmap is 0
 * elsewhere).  Optional outputs (NULL skips): almB (nsim, lmax+1, lmax+1) the
 * B alm of the masked maps, Btpl (npix, nsim) the template map alm2t(B'),
 * coef (2, nsim) the fit (a0, a1) per map.  Needs pol != 0.  Allocates its
 * temporaries (4 alm sets, 2 map sets; host calls 4 more map sets) for the
 * duration of the call and returns when done. */
int fsht_fix_eb(fsht_handle* h, double* Q, double* U, const double* mask, double* Bmap, double* almB, double* Btpl,
                double* coef, int niter, int loc);

/* fsht_fix_eb on device arrays (as loc = FSHT_DEVICE) without any device
 * allocation.  No counterpart in the reference; it is for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops).  Its
 * temporaries (3 alm sets and 2 map sets, plus almB / Btpl / coef when they
 * are NULL, each rounded up to 256 bytes) come from a static device array of
 * 128 MiB that every device holds from library load to process end; larger
 * configurations return FSHT_ERR_ARG.  niter > 0 needs a handle created with
 * niter > 0.  Asynchronous on the handle's stream.  All calls on one GPU share
 * the array (it keeps the last call's contents; what must start at zero is
 * cleared by the call), so calls that overlap on one GPU race. */
int fsht_fix_eb_noalloc(fsht_handle* h, double* Q, double* U, const double* mask, double* Bmap, double* almB,
                        double* Btpl, double* coef, int niter);

/* Page-locked host memory for map / alm arrays: host-pointer calls move
 * such arrays by DMA directly (other host arrays go through the handle's
 * pinned staging buffers).  Replaces the numba.cuda.pinned_array outputs of
 * SHT.py (SHT.py:80-84, 110-114). */
int fsht_host_alloc(unsigned long long bytes, void** out);
int fsht_host_free(void* p);

/* Per-kernel-class timing with HIP events on the handle's stream.
 * fsht_profile(h, 1) starts recording (and clears earlier records); classes:
 * 0 ring FFT forward, 1 ring FFT inverse, 2 transposes, 3 Legendre analysis,
 * 4 Legendre synthesis, 5 alias fold / Jacobi residual, 6 spectrum gather
 * into the m-major analysis operand, 7 operand slicing (mixed precision),
 * 8 spin-2 Legendre table generation (mixed precision).  fsht_profile_read
 * fills out[2c] = launches, out[2c+1] = total ms (c < 9). */
int fsht_profile(fsht_handle* h, int enable);
int fsht_profile_read(fsht_handle* h, double* out, int n);

/* alm utilities on device arrays (stream: hipStream_t or NULL). */
/* out = in*fac(l) + shift(l), with an extra factor m0 on m = 0.
 * Replaces norm_alm1 (src/norm_alm.f90:14-33). */
int fsht_norm_alm(const double* alm_in, double* alm_out, const double* fac, const double* shift, double m0,
                  int lmax, int nsim, void* stream);
/* alm *= win(l) in place.  Replaces filt_alm1 (src/filt_alm.f90:16-36). */
int fsht_filt_alm(double* alm, const double* win, int lmax, int nsim, void* stream);
/* Square packing -> healpy m-major order: alm_hp (2, (lmax+1)(lmax+2)/2, nsim)
 * Fortran order (Re, Im interleaved).  Replaces convert_alm_healpy1
 * (src/convert_alm.f90:34-56) used by SHT.convert_alm_healpy (SHT.py:204-210).
 * loc = FSHT_DEVICE: device arrays, asynchronous on `stream`; loc = FSHT_HOST:
 * host arrays, staged through device buffers, returns with the result. */
int fsht_convert_alm_healpy(const double* alm, double* alm_hp, int lmax, int nsim, int loc, void* stream);
/* Inverse of the above.  Replaces convert_alm_healpy_inv1 (src/convert_alm.f90:87-107). */
int fsht_convert_alm_healpy_inv(const double* alm_hp, double* alm, int lmax, int nsim, int loc, void* stream);

/* Angular power spectra of packed alm (device arrays): T only (almE = almB =
 * NULL): cl (nsim, lmax+1); or all of TT EE BB TE TB EB: cl (nsim, lmax+1, 6),
 * Fortran order.  cl(k, l, c) += sum_m fac_m * (Re x Re y + Im x Im y)/(2l+1)
 * for l >= lmin (fac_0 = 1, fac_m = 2); like the reference it accumulates into
 * cl, so pass zeros for a plain spectrum.  Replaces alm2cl_T1 / alm2cl_TEB1
 * (src/alm2cl_TEB.f90:22-110). */
int fsht_alm2cl(const double* almT, const double* almE, const double* almB, double* cl, int lmin, int lmax, int nsim,
                void* stream);
/* Seeded Gaussian alm with spectrum cl (device arrays): T only (almE = almB =
 * NULL, cl has lmax+1 entries) or T, E, B with cl (lmax+1, 6) = TT EE BB TE TB
 * EB; elements with l < lmin are 0.  Replaces cl2alm_T1 / cl2alm_TEB1
 * (src/cl2alm_TEB.f90:25-170) with the same per-l linear combination, but the
 * random stream is Philox4x32-10 + Box-Muller keyed by seed (MKL VSL's stream
 * is not reproduced); each element depends only on (seed, sim, element). */
int fsht_cl2alm(const double* cl, double* almT, double* almE, double* almB, int lmin, int lmax, int nsim,
                unsigned long long seed, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* FASTSHT_H */
