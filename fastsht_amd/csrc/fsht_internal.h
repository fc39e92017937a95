// Internal declarations shared by the fastsht_amd HIP sources (not part of the C ABI).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

namespace fsht {

// -------------------------------------------------------------------------
// Geometry of one north ring i in [0, 2*nside) and its south mirror.
// -------------------------------------------------------------------------
struct RingGeo {
  long long i1N;  // first pixel of north ring i
  long long i1S;  // first pixel of its mirror nring-1-i (== i1N for the equator)
  int nfft;       // ring length
  int pad_;
  double z;       // cos(theta)
  double s;       // sin(theta)
};

// One ring-FFT job: a north ring and its mirror as one complex transform of
// length n (half = 0), or, for rings whose pair transform would not fit in
// LDS, one of the two real rings alone as a complex transform of length n/2
// (half = 1: north ring, 2: south ring).
struct FftJob {
  int ring;      // north ring index
  int n;         // ring length
  int M;         // transform length used in LDS (the transform length if a power of two, else Bluestein size)
  int blue;      // index into the Bluestein tables, -1 for direct power-of-two transforms
  int half;      // 0: N/S pair; 1 or 2: north or south ring alone (length-n/2 transform)
};

struct BlueInfo {
  long long chirp_off;  // offset of w_j = exp(-i pi j^2 / L), j < L (L = Bluestein length),
                        // then exp(-i pi k / L), k = 0..L (the real-spectrum split roots)
  long long filt_off;   // offset of FFT_M(b)/M (M entries, bit-reversed order)
};

// Work item of the Legendre kernels.
struct LegItem {
  int m;
  int par;     // parity / symmetry class (0: l+m even, 1: odd)
  int blk;     // analysis: l-block; synthesis: ring block
  int colblk;  // analysis: block of kAnaCols columns; synthesis: first 16-column tile
  int r0;      // analysis: first ring (from i0(m), a multiple of 4) with a kept P_lm in the l-block
};

// Tunables (also used by the host to size work lists).
constexpr int kWG = 256;          // threads per Legendre workgroup
constexpr int kAnaRings0 = 256;   // analysis ring sub-block (spin 0)
constexpr int kAnaRings2 = 128;   // analysis ring sub-block (spin 2, x2 functions)
constexpr int kAnaLB = 256;       // analysis l-block (l steps)
constexpr int kAnaCols = 64;      // analysis columns per workgroup (4 waves x 16)
constexpr int kSynRings = 256;    // synthesis rings per workgroup (4 waves x 64)
constexpr int kSynCols = 64;      // synthesis columns per wave (4 tiles of 16; spin 0 batches end with a 2-tile block)
constexpr int kSegL = 256;        // synthesis: recurrence coefficients staged per wave (l values)
// P_lm (and the spin-2 functions built from it) below 2^-100 is dropped:
// the recurrence of a ring starts at its first l with |P_lm| >= 2^-100.  A
// dropped term changes a sum of O(1) terms by < 2^-100 * 10^8 (10^8 bounds
// the largest spin-2 factor, 2/sin^2 at the polar ring), far below FP64 rounding.
constexpr int kKeepExp = -99;     // seeds: a value p*2^e is kept iff frexp-exponent(p)+e >= this

// ---------------------------------------------------------------------------
// Mixed precision: the Legendre contractions as sums of int8 x int8 products
// of fixed-point digits (v_mfma_i32_32x32x32_i8, exact int32 accumulation).
// A value v of a row or column with exponent e is x = v * 2^-e in
// [-127/128, 127/128], written as K balanced base-128 digits
// x = sum_i d_i 2^-7(i+1) (|d_0| <= 127, |d_i| <= 64 after).  Products keep
// the digit pairs i + j < K.  Every row of the left operand and every column
// of the right one has one exponent over the whole contraction, so the int32
// digit sums are converted to FP64 once, after the last step.  Operand tiles
// are 1 KB blocks in the MFMA operand order: lane L = idx + 32 g holds the 16
// bytes of contraction index 16 g .. 16 g + 15 (idx: l or ring row of the
// tile, or column).  Per (m, parity class) the layouts are
//   TA  (P for the analysis):  [l-tile][ring step][slice]  lanes: l, 16 rings each
//   TS  (P for the synthesis): [ring tile][l step][slice]  lanes: ring, 16 l each
//   FQ  (analysis operand):    [col tile][ring step][slice]
//   AQ  (synthesis operand):   [col tile][l step][slice]
// with tiles and steps of 32 (ring rows from i0(m); l as class values
// j = (l - m - class) / 2).  Exponents: TA per l and TS per ring row (32 per
// tile, stored in accumulator-row order), FQ and AQ per column.
// ---------------------------------------------------------------------------
constexpr int kMixK = 6;           // slices stored in the P tables
constexpr int kMixZeroE = -30000;  // exponent of an all-zero row or column
// Exponent e of a row or column with max |v| = mx: |v| * 2^-e <= 127/128.
__device__ inline int mix_exp(double mx) { return mx > 0.0 ? ilogb(mx * (128.0 / 127.0)) + 1 : kMixZeroE; }
struct MixOff {
  long long ta, ts;    // byte offsets into the P tables
  long long fq, aq;    // byte offsets into the sliced operands (room for kMixK slices)
  long long tae, tse;  // offsets (ints) of the P-table exponents
  long long fqe, aqe;  // offsets (ints) of the operand exponents
  int nrs;             // ring steps (= ring tiles) of 32 rows: ceil(rows / 32)
  int nlt;             // l steps (= l tiles) of 32 class values
  int fa, fs;          // offsets into the first-nonzero-step tables
};

struct Handle;

// Kernel launchers (k_*.hip). All enqueue on h->stream.
hipError_t launch_seeds(Handle* h);
hipError_t build_bluestein(Handle* h);
// maps [map0, map0 + nmap) of the comp*nsim + sim index space (nmap <= 0: all)
hipError_t launch_ring_fft(Handle* h, const double* const* maps_comp, double* spec, int inverse,
                           double* const* out_comp, int map0 = 0, int nmap = 0);
// map-major ring spectra -> m-major analysis operand.  fexp (mixed precision,
// else NULL): per (m, class, column) the max over rings of the exponent of |F|
// is folded in with atomicMax (reset to kMixZeroE first)
hipError_t launch_gather(Handle* h, const double* spec, double* F, int* fexp = nullptr);
// residual analysis operand F0 - n * gather(fold(G)) straight from the
// synthesis buffer G and the first operand F0: aliases m' <= m (full = 0,
// immediate update) or all aliases (full = 1, Jacobi)
hipError_t launch_gather_prefix(Handle* h, const double* G, const double* F0, double* F, int full,
                                int* fexp = nullptr);
hipError_t launch_analysis(Handle* h, const double* F, double* alm0, double* alm1, double beta,
                           int want0, int want1);
hipError_t launch_synthesis(Handle* h, const double* alm0, const double* alm1, double* G);
// m-major synthesis output -> map-major ring spectra
hipError_t launch_fold(Handle* h, const double* G, double* spec);
// XCD-ordered upload of a work list (setup.hip): items of one m stay on one XCD,
// m-groups spread over the XCDs by the estimated cost of each item
hipError_t upload_items(const std::vector<LegItem>& v, const std::vector<double>& cost, LegItem** d, int* n);
// mixed precision (k_mixed.hip)
hipError_t mixed_setup(Handle* h);        // shared state, and the P tables of a temperature-only handle
hipError_t mixed_ensure(Handle* h, int spin);  // the state a spin-0 or spin-2 transform needs (set-up on first use)
// Device bytes the mixed-precision state would take, planned on the host:
// out[0] the shared sliced operands, out[1] the spin-0 tables, out[2] the
// spin-2 exponents and range buffers.
void mixed_planned_bytes(const Handle* h, long long out[3]);
void mixed_free(Handle* h);
hipError_t launch_fexp_reset(Handle* h);                          // FQe <- kMixZeroE
hipError_t launch_quant_F(Handle* h, const double* F, int K);      // m-major FP64 operand -> FQ (FQe set)
// packed alm (alm1: the second component, spin 2) -> AQ; have_exp: AQe
// already holds the columns' exponents (set by the analysis that wrote the
// alm with exps = 1), else they are computed here
hipError_t launch_quant_A(Handle* h, const double* alm0, const double* alm1, int K, int have_exp);
// Slice-GEMM passes run over ranges of orders: one range for spin 0; for
// spin 2 each range's F1/F2 tables are generated (launch_ftab) just before
// its contractions.  An analysis that records its result's column exponents
// (exps = 1) needs launch_aqe_reset first.
int mixed_nrange(const Handle* h);
hipError_t launch_ftab(Handle* h, int range, int synthesis, int K);
hipError_t launch_aqe_reset(Handle* h);
hipError_t launch_analysis_q(Handle* h, int range, double* alm0, double* alm1, double beta, int K, int exps);
hipError_t launch_synthesis_q(Handle* h, int range, double* G, int K);
// map-domain helpers (k_maps.hip)
hipError_t launch_mask_maps(double* maps, const double* mask, long long npix, int nmaps, hipStream_t s);
hipError_t launch_zero(double* p, long long n, hipStream_t s);
hipError_t launch_linfit_resi(const double* x, const double* y, const double* mask, double* resi, double* coef,
                              long long npix, int nmaps, hipStream_t s);

// -------------------------------------------------------------------------
// Handle: one transform configuration with its resident tables and workspace.
// -------------------------------------------------------------------------
struct Handle {
  int nside = 0, lmax = 0, nsim = 0, niter = 0, pol = 0;
  int nring = 0, nhr = 0, nnorth = 0;  // nnorth = 2*nside rings incl. equator
  long long npix = 0;
  int nsimp = 0;  // capacity rounded up to 16
  int nsim_cap = 0;  // maps the handle was created for (nsim <= nsim_cap: fsht_set_batch)
  double pa = 0;
  hipStream_t stream = nullptr;      // stream all work is enqueued on
  hipStream_t own_stream = nullptr;  // stream created (and destroyed) by the handle
  int device = 0;
  void* scratch = nullptr;  // this device's copy of the static scratch array (fsht_fix_eb_noalloc)
  std::string err;

  // host tables
  std::vector<int> i0_ref;      // ring_tab per m (0..lmax): the reference's ring-skip table
  std::vector<int> i0;          // first ring used per m: ring_tab raised past rings whose P_lm
                                // is dropped for every l <= lmax (non-decreasing in m)
  std::vector<long long> aboff; // offset of order m in the (a,b) table
  std::vector<long long> seedoff;
  std::vector<long long> synoff;  // offset of order m in the per-(m, ring) synthesis seeds
  std::vector<int> nblk;        // analysis l-blocks per m
  std::vector<long long> goff;  // offset (in doubles) of order m in the m-major synthesis buffer

  // device tables
  RingGeo* d_geo = nullptr;
  int* d_i0 = nullptr;
  double2* d_ab = nullptr;      // (a_l, b_l) per (m, l)
  double2* d_spin = nullptr;    // (lam_l, normal_l) per (m, l)
  long long* d_aboff = nullptr;
  double2* d_tw = nullptr;      // (cos m phi0, sin m phi0) per (m, north ring)
  int* d_slot = nullptr;        // relative PERM slot*4 + class per (m, north ring)
  long long* d_seedoff = nullptr;
  // per (m, ring, l-block): first kept l in the block, st, and the recurrence
  // state at the 32-step chunk start ls <= st (l - m = 0 mod 32)
  int* d_seed_st = nullptr;
  double2* d_seed_v = nullptr;  // (P_ls-2, P_ls-1)
  double* d_seed_p = nullptr;   // P_ls
  long long* d_synoff = nullptr;
  int* d_syn_st = nullptr;      // first l whose P_lm is kept, per (m, ring): the synthesis start
  double2* d_syn_v = nullptr;   // (P_ls-2, P_ls-1) at the chunk start ls <= that l
  long long* d_goff = nullptr;
  LegItem* d_ana_items[2][2] = {{nullptr, nullptr}, {nullptr, nullptr}};  // [spin 0/2][full / 2-tile column blocks]
  LegItem* d_syn_items[2] = {nullptr, nullptr};
  int n_ana_items[2][2] = {{0, 0}, {0, 0}}, n_syn_items[2] = {0, 0};

  // FFT tables
  std::vector<FftJob> fft_jobs;          // host copy, grouped by class
  std::vector<int> fft_class_begin;      // start index of each LDS class in fft_jobs
  std::vector<int> fft_class_M;          // max M of each class
  FftJob* d_fft_jobs = nullptr;
  BlueInfo* d_blue = nullptr;
  double2* d_chirp = nullptr;
  double2* d_filt = nullptr;
  double2* d_twid = nullptr;  // exp(-2 pi i k / Mmax), k < Mmax
  int Mmax = 0;
  std::vector<BlueInfo> blue_host;
  std::vector<int> blue_n, blue_M;

  // workspace (device)
  int ncomp_ws = 0;            // components the workspace is sized for (1 or 2)
  double* ws_spec = nullptr;   // ncomp*nsim maps of npix (sim-major)
  double* ws_F0 = nullptr;     // m-major analysis operand of the input maps (kept for the refinement)
  double* ws_G = nullptr;      // m-major synthesis output
  double* ws_F = nullptr;      // m-major residual analysis operand of the refinement sweeps
  long long G_size = 0;
  double* ws_alm[2] = {nullptr, nullptr};  // device alm (nsim,(lmax+1)^2) for host-pointer calls
  double* ws_map[2] = {nullptr, nullptr};  // device maps for host-pointer calls
  int have_dev_alm = 0;  // result of the last forward transform kept on device

  // pinned staging of host arrays (two buffers, allocated on first host call)
  double* stage[2] = {nullptr, nullptr};
  hipEvent_t stage_ev[2] = {nullptr, nullptr};
  size_t stage_bytes = 0;

  // current call's spin (0 or 2)
  int spin = 0;

  // mixed precision: int8-slice Legendre tables, sliced operands, work lists
  // and the slice counts per pass
  int precision = 0;  // 0: FP64, 1: mixed
  int noalloc = 0;    // inside an allocation-free call: no state is set up on first use
  int mix_ready = 0;   // offsets, sliced operands and table work items set up
  int mix0_ready = 0;  // spin 0: P tables and work lists
  int mix2_ready = 0;  // spin 2: F1/F2 exponents, m ranges, chunk buffers and work lists
  std::vector<MixOff> mix_off;  // per (m, class)
  MixOff* d_mix_off = nullptr;
  signed char *d_TA = nullptr, *d_TS = nullptr, *d_FQ = nullptr, *d_AQ = nullptr;
  int fq_alias = 0;  // d_FQ lives in ws_spec
  int *d_TAe = nullptr, *d_TSe = nullptr, *d_FQe = nullptr, *d_AQe = nullptr;
  int *d_firstA = nullptr, *d_firstS = nullptr;
  int2* d_pitems = nullptr;        // (m, 256-row block) items of the table kernels, m ascending
  std::vector<int> pitem_off;      // first item of each m (lmax + 2 entries)
  long long mix_ta_total = 0, mix_tae = 0, mix_tse = 0;  // table bytes (six slices) and exponent counts
  int mix_fa = 0, mix_fs = 0;      // entries of the first-nonzero-step tables
  // slice-GEMM work lists: [0] full blocks of 4 column tiles (4 row tiles
  // each), [1] the remaining column tiles one by one (8 row tiles each), [2]
  // the remaining column tiles as one padded block of 4 (for six slices,
  // whose 8-row-tile stages would not fit in LDS)
  LegItem* d_anaq_items[3] = {nullptr, nullptr, nullptr};
  LegItem* d_synq_items[3] = {nullptr, nullptr, nullptr};
  int n_anaq_items[3] = {0, 0, 0}, n_synq_items[3] = {0, 0, 0};
  // spin 2: F1 and F2 exponents / first nonzero steps (resident), the tables
  // of one m range in one operand order (generated per pass), the ranges and
  // the work lists per (function, range, kind)
  int *d_TAe2[2] = {nullptr, nullptr}, *d_TSe2[2] = {nullptr, nullptr};
  int *d_firstA2[2] = {nullptr, nullptr}, *d_firstS2[2] = {nullptr, nullptr};
  signed char* d_T2[2] = {nullptr, nullptr};
  long long t2_cap = 0;
  std::vector<int> mix2_chunk;  // range c holds the orders [mix2_chunk[c], mix2_chunk[c + 1])
  std::vector<LegItem*> d_anaq2, d_synq2;  // index ((f * nchunk) + c) * 3 + kind
  std::vector<int> n_anaq2, n_synq2;
  int t2_lay = -1, t2_ks = 0;  // operand order and slices the chunk buffers hold (single-range plans)
  long long mix_bytes = 0;  // device memory held by the mixed-precision tables and operands
  int mix_nrb_max = 0;  // 256-row blocks of the longest order (grid of the operand slicer)
  // slices per pass: the first analysis and a residual synthesis whose error
  // the remaining Jacobi iterations of t2alm damp at least mix_damp_min
  // times, and a later analysis (of a residual) damped at least
  // mix_damp_min_a times, use mix_k_lo, the others mix_k_hi
  // (set from lmax / nside: the stronger the aliasing, the weaker the
  // damping); alm2t and an unrefined t2alm use mix_k_out
  int mix_k_lo = 4, mix_k_hi = 5, mix_damp_min = 2, mix_k_out = 6, mix_damp_min_a = 1;
  // the slices of every pass per spin (0: temperature, 1: spin 2): [kind][d]
  // with kind 0 the first analysis, 1 a residual synthesis, 2 a residual
  // analysis, and d = min(2, later Jacobi iterations that damp the pass's
  // error); mix_out: alm2t / eb2qu and an unrefined analysis
  int mix_lad[2][3][3] = {{{6, 5, 4}, {5, 5, 4}, {5, 4, 4}}, {{6, 5, 4}, {6, 5, 4}, {5, 4, 4}}};
  int mix_out[2] = {6, 6};

  // optional per-kernel-class timing with HIP events on `stream`
  int profiling = 0;
  struct Timed {
    int cls;
    hipEvent_t a, b;
  };
  std::vector<Timed> timed;
  std::vector<hipEvent_t> event_pool;
};

// Spin-2 functions from P_l and P_{l-1} (src/get_plm_recursive.f90:111-124).
// (f = 0: F1, 1: F2; sp = (lambda_l, normal_l); bw = cos/sin^2, two_on_s2 = 2/sin^2,
// cw = m/sin^2)
__device__ inline double spin_f(int f, double P, double Pm1, double2 sp, int l, int m, double x, double bw,
                                double two_on_s2, double cw) {
  const double lam = sp.x, nl = sp.y;
  const double dl = (double)l;
  if (f == 0) {
    const double lm2 = dl - (double)m * (double)m;
    return (Pm1 * lam * bw - P * (two_on_s2 * lm2 + dl * (dl - 1.0))) * nl;
  }
  return (Pm1 * lam - P * (2.0 * x * (dl - 1.0))) * cw * nl;
}

// kernel classes reported by fsht_profile_read
// Batches smaller than the handle's capacity: the m-major buffers keep nsimp
// columns (the capacity rounded to 16) per Re/Im and component segment, of
// which the first nsim hold maps.  True if a column of [c0, c1) holds a map.
__device__ inline bool cols_active(int c0, int c1, int nsim, int nsimp) {
  for (int c = c0; c < c1;) {
    const int seg = c / nsimp, s = c - seg * nsimp;
    if (s < nsim) return true;
    c = (seg + 1) * nsimp;
  }
  return false;
}

enum KernelClass {
  KC_FFT_FWD = 0,
  KC_FFT_INV,
  KC_TRANSPOSE,  // no launches (the gathers and folds read and write the FFT's layout)
  KC_ANALYSIS,
  KC_SYNTHESIS,
  KC_FOLD,
  KC_GATHER,
  KC_SLICE,  // mixed precision: operand slicing (k_quant_F / k_quant_A)
  KC_TABLES, // mixed precision: spin-2 F1/F2 tables of a range of orders (k_ftab)
  KC_COUNT
};

}  // namespace fsht
