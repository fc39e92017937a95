// C ABI of libfastsht (declared in include/fastsht.h).
//
// One handle = one (nside, lmax, nsim, niter, pol) transform configuration
// with its resident tables and workspace: the state the reference keeps in
// process-global Fortran module variables (src/sht_data_module.f90:10-109),
// made per-instance.  Every entry point returns 0 on success or a negative
// status; fsht_last_error() gives the message.
#include <cstdio>
#include <cstring>
#include <string>

#include <algorithm>
#include <mutex>
#include <vector>
#include <thread>

#include "fastsht.h"
#include "fsht_internal.h"

namespace fsht {
hipError_t setup_tables(Handle* h);
hipError_t launch_norm_alm(const double*, double*, const double*, const double*, double, int, int, hipStream_t);
hipError_t launch_filt_alm(double*, const double*, int, int, hipStream_t);
hipError_t launch_convert_healpy(const double*, double*, int, int, int, hipStream_t);
hipError_t launch_alm2cl(const double*, const double*, const double*, double*, int, int, int, int, hipStream_t);
hipError_t launch_cl2alm(const double*, double*, double*, double*, int, int, int, int, unsigned long long, hipStream_t);
void synth_offsets(Handle* h, int ncomp);
}  // namespace fsht

using fsht::Handle;

struct fsht_handle {
  Handle h;
};

static thread_local std::string g_err;

// Scratch of the allocation-free entry point (fsht_fix_eb_noalloc): a
// file-scope device array of kScratchBytes, one copy per device that loads
// the library, held until the process ends.  Host code never uses the array's
// name as an address: its device address is fetched once per device with
// hipGetSymbolAddress (with that device current) and kept in g_scratch.
constexpr size_t kScratchBytes = 128u << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];
constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch[kMaxDevices] = {};

// Device address of the current device's scratch (device `dev` is current).
static hipError_t scratch_address(int dev, void** out) {
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch[dev] = p;
  }
  *out = g_scratch[dev];
  return hipSuccess;
}

static int fail(Handle* h, int code, const std::string& msg) {
  if (h) h->err = msg;
  g_err = msg;
  return code;
}

#define HIPCHK(h, expr)                                                                              \
  do {                                                                                               \
    hipError_t _e = (expr);                                                                          \
    if (_e != hipSuccess)                                                                            \
      return fail((h), FSHT_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e));             \
  } while (0)

static void free_all(Handle* h) {
  fsht::mixed_free(h);
  void* ptrs[] = {h->d_geo, h->d_i0, h->d_ab, h->d_spin, h->d_aboff, h->d_tw, h->d_slot, h->d_seedoff,
                  h->d_seed_st, h->d_seed_v, h->d_seed_p, h->d_synoff, h->d_syn_st, h->d_syn_v, h->d_goff, h->d_ana_items[0][0], h->d_ana_items[0][1],
                  h->d_ana_items[1][0], h->d_ana_items[1][1], h->d_syn_items[0], h->d_syn_items[1], h->d_fft_jobs, h->d_blue, h->d_chirp, h->d_filt,
                  h->d_twid, h->ws_spec, h->ws_F0, h->ws_G, h->ws_F, h->ws_alm[0], h->ws_alm[1]};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  for (int b = 0; b < 2; ++b) {
    if (h->stage[b]) (void)hipHostFree(h->stage[b]);
    if (h->stage_ev[b]) (void)hipEventDestroy(h->stage_ev[b]);
  }
}

static bool is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }

extern "C" {

const char* fsht_version(void) { return "fastsht_amd 0.2 (gfx950, FP64 MFMA; mixed: int8-slice MFMA)"; }

const char* fsht_last_error(const fsht_handle* h) {
  if (h && !h->h.err.empty()) return h->h.err.c_str();
  return g_err.c_str();
}

int fsht_create(int nside, int lmax, int nsim, int niter, int pol, fsht_handle** out) {
  if (!out) return fail(nullptr, FSHT_ERR_ARG, "out handle pointer is NULL");
  *out = nullptr;
  if (!is_pow2(nside) || nside > 2048)
    return fail(nullptr, FSHT_ERR_ARG, "nside must be a power of two in [1, 2048]");
  if (lmax < 0 || lmax > 4 * nside)
    return fail(nullptr, FSHT_ERR_ARG, "lmax must be in [0, 4*nside]");
  if (nsim < 1) return fail(nullptr, FSHT_ERR_ARG, "nsim must be >= 1");
  // the Legendre kernels address one alm column (nsim x (lmax+1) doubles) with 32-bit byte offsets
  if ((long long)nsim * (lmax + 1) >= (1LL << 28))
    return fail(nullptr, FSHT_ERR_ARG, "nsim*(lmax+1) must be below 2^28");
  // ... and one order's folded spectrum (2*nside rings x 4 column blocks of nsim
  // rounded to 16, spin 2) likewise
  if (16LL * nside * 4 * (((long long)nsim + 15) / 16 * 16) >= (1LL << 32))
    return fail(nullptr, FSHT_ERR_ARG, "nside*nsim too large (one order's spectrum must stay below 4 GiB)");
  if (niter < 0) return fail(nullptr, FSHT_ERR_ARG, "niter must be >= 0");
  fsht_handle* H = new fsht_handle();
  Handle* h = &H->h;
  h->nside = nside;
  h->lmax = lmax;
  h->nsim = nsim;
  h->nsim_cap = nsim;
  h->niter = niter;
  h->pol = pol ? 1 : 0;
  h->nring = 4 * nside - 1;
  h->nhr = 2 * nside - 1;
  h->nnorth = 2 * nside;
  h->npix = 12LL * nside * nside;
  h->nsimp = (nsim + 15) / 16 * 16;
  h->pa = 4.0 * 3.141592653589793238462643383279502884 / (double)h->npix;
  hipError_t e = hipGetDevice(&h->device);
  if (e == hipSuccess) e = scratch_address(h->device, &h->scratch);
  if (e == hipSuccess) e = hipStreamCreate(&h->own_stream);
  h->stream = h->own_stream;
  if (e == hipSuccess) e = fsht::setup_tables(h);
  if (e == hipSuccess) e = fsht::build_bluestein(h);
  if (e != hipSuccess) {
    std::string msg = h->err.empty() ? std::string("setup: ") + hipGetErrorString(e) : h->err;
    free_all(h);
    if (h->own_stream) (void)hipStreamDestroy(h->own_stream);
    delete H;
    return fail(nullptr, e == hipErrorInvalidValue ? FSHT_ERR_ARG : FSHT_ERR_HIP, msg);
  }
  // workspace
  const int ncomp = h->pol ? 2 : 1;
  h->ncomp_ws = ncomp;
  fsht::synth_offsets(h, 2);
  const long long nspec = (long long)ncomp * nsim * h->npix;
  const long long nG = h->G_size / (ncomp == 2 ? 1 : 2);
  auto alloc = [&](double** p, long long n) -> hipError_t { return hipMalloc((void**)p, (size_t)n * sizeof(double)); };
  e = alloc(&h->ws_spec, nspec);
  if (e == hipSuccess) e = alloc(&h->ws_F0, nG);
  if (e == hipSuccess) e = alloc(&h->ws_G, nG);
  if (e == hipSuccess && niter > 0) e = alloc(&h->ws_F, nG);
  if (e != hipSuccess) {
    free_all(h);
    (void)hipStreamDestroy(h->own_stream);
    delete H;
    return fail(nullptr, FSHT_ERR_NOMEM, std::string("workspace allocation: ") + hipGetErrorString(e));
  }
  *out = H;
  return FSHT_OK;
}

int fsht_destroy(fsht_handle* H) {
  if (!H) return FSHT_OK;
  Handle* h = &H->h;
  (void)hipStreamSynchronize(h->stream);
  free_all(h);
  for (auto& t : h->timed) {
    (void)hipEventDestroy(t.a);
    (void)hipEventDestroy(t.b);
  }
  for (auto e : h->event_pool) (void)hipEventDestroy(e);
  (void)hipStreamDestroy(h->own_stream);
  delete H;
  return FSHT_OK;
}

int fsht_synchronize(fsht_handle* H) {
  if (!H) return fail(nullptr, FSHT_ERR_ARG, "NULL handle");
  HIPCHK(&H->h, hipStreamSynchronize(H->h.stream));
  return FSHT_OK;
}

int fsht_get_stream(fsht_handle* H, void** stream) {
  if (!H || !stream) return fail(nullptr, FSHT_ERR_ARG, "NULL argument");
  *stream = (void*)H->h.stream;
  return FSHT_OK;
}

int fsht_set_stream(fsht_handle* H, void* stream) {
  if (!H) return fail(nullptr, FSHT_ERR_ARG, "NULL handle");
  Handle* h = &H->h;
  HIPCHK(h, hipStreamSynchronize(h->stream));
  h->stream = (hipStream_t)stream;  // NULL: the null stream, ordered with every blocking stream
  return FSHT_OK;
}

int fsht_set_batch(fsht_handle* H, int nsim) {
  if (!H) return fail(nullptr, FSHT_ERR_ARG, "NULL handle");
  Handle* h = &H->h;
  if (nsim < 1 || nsim > h->nsim_cap)
    return fail(h, FSHT_ERR_ARG, "batch must hold 1.." + std::to_string(h->nsim_cap) + " maps");
  h->nsim = nsim;
  h->have_dev_alm = 0;  // kept alms were laid out for the previous batch
  return FSHT_OK;
}

int fsht_info(const fsht_handle* H, long long* out, int n) {
  if (!H || !out) return fail(nullptr, FSHT_ERR_ARG, "NULL argument");
  const Handle* h = &H->h;
  long long v[39] = {h->nside,    h->lmax,     h->nsim,     h->niter,        h->pol,
                     h->npix,     h->nsimp,    h->G_size,   h->precision,    h->mix_bytes,
                     h->mix_k_lo, h->mix_k_hi, h->mix_damp_min, h->mix_k_out, h->mix_damp_min_a,
                     h->nsim_cap};
  for (int sp = 0; sp < 2; ++sp) {
    for (int k = 0; k < 9; ++k) v[16 + 10 * sp + k] = h->mix_lad[sp][k / 3][k % 3];
    v[25 + 10 * sp] = h->mix_out[sp];
  }
  fsht::mixed_planned_bytes(h, v + 36);
  for (int i = 0; i < n && i < 39; ++i) out[i] = v[i];
  return FSHT_OK;
}

int fsht_ring_tab(const fsht_handle* H, int* out) {
  if (!H || !out) return fail(nullptr, FSHT_ERR_ARG, "NULL argument");
  std::memcpy(out, H->h.i0_ref.data(), H->h.i0_ref.size() * sizeof(int));
  return FSHT_OK;
}

int fsht_set_precision(fsht_handle* H, int precision) {
  if (!H) return fail(nullptr, FSHT_ERR_ARG, "NULL handle");
  Handle* h = &H->h;
  if (precision != FSHT_PREC_FP64 && precision != FSHT_PREC_MIXED)
    return fail(h, FSHT_ERR_ARG, "precision must be FSHT_PREC_FP64 or FSHT_PREC_MIXED");
  if (precision == FSHT_PREC_MIXED && !h->mix_ready) {
    HIPCHK(h, hipStreamSynchronize(h->stream));
    const hipError_t e = fsht::mixed_setup(h);
    if (e == hipErrorInvalidValue) return fail(h, FSHT_ERR_ARG, "mixed precision needs lmax - m <= 4607");
    if (e == hipErrorOutOfMemory)
      return fail(h, FSHT_ERR_NOMEM, "mixed precision: the int8 Legendre tables and sliced operands do not fit");
    if (e != hipSuccess) return fail(h, FSHT_ERR_HIP, std::string("mixed precision set-up: ") + hipGetErrorString(e));
  }
  // slice counts: 4 for the first analysis and residual syntheses damped at
  // least twice and for the later analyses (of residuals) damped at least
  // once, 5 otherwise, while lmax <= 2.5 nside (Jacobi contracts by ~0.1 per
  // iteration there); stronger
  // aliasing contracts less, so 5 everywhere up to lmax = 3.5 nside and 6
  // beyond; 6 everywhere for nside <= 16
  if (h->nside <= 16) {  // coarse grids contract least; their passes are cheap
    h->mix_k_lo = h->mix_k_hi = 6;
  } else if (2 * h->lmax <= 5 * h->nside) {
    h->mix_k_lo = 4;
    h->mix_k_hi = 5;
  } else if (2 * h->lmax <= 7 * h->nside) {
    h->mix_k_lo = h->mix_k_hi = 5;
  } else {
    h->mix_k_lo = h->mix_k_hi = 6;
  }
  h->mix_damp_min = 2;    // residual syntheses
  h->mix_damp_min_a = 1;  // analyses (after the first, they act on residuals)
  h->mix_k_out = 6;
  // the same rule as a table per (pass kind, damping).  Spin 2 adds one
  // slice to the residual synthesis no later iteration damps: with five
  // there, qu2eb(niter >= 1) sits at 1.0-1.5e-10 against the FP64 result at
  // nside 64-1024 (lmax = 2 nside - 1); with six, at <= 1.5e-11
  for (int d = 0; d < 3; ++d) {
    h->mix_lad[0][0][d] = d == 0 ? h->mix_k_out : (d >= h->mix_damp_min ? h->mix_k_lo : h->mix_k_hi);
    h->mix_lad[0][1][d] = d >= h->mix_damp_min ? h->mix_k_lo : h->mix_k_hi;
    h->mix_lad[0][2][d] = d >= h->mix_damp_min_a ? h->mix_k_lo : h->mix_k_hi;
    for (int k = 0; k < 3; ++k) h->mix_lad[1][k][d] = h->mix_lad[0][k][d];
  }
  h->mix_lad[1][1][0] = std::min(6, h->mix_lad[0][1][0] + 1);
  h->mix_out[0] = h->mix_out[1] = h->mix_k_out;
  h->precision = precision;
  return FSHT_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Optional per-kernel-class timing: HIP events recorded on the handle's stream
// around each launch (fsht_profile / fsht_profile_read).
// ---------------------------------------------------------------------------
static hipEvent_t get_event(Handle* h) {
  if (!h->event_pool.empty()) {
    hipEvent_t e = h->event_pool.back();
    h->event_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  (void)hipEventCreate(&e);
  return e;
}

template <class F>
static hipError_t timed(Handle* h, int cls, F&& f) {
  if (!h->profiling) return f();
  hipEvent_t a = get_event(h), b = get_event(h);
  hipError_t e = hipEventRecord(a, h->stream);
  if (e != hipSuccess) return e;
  e = f();
  if (e != hipSuccess) return e;
  e = hipEventRecord(b, h->stream);
  h->timed.push_back(Handle::Timed{cls, a, b});
  return e;
}

// ---------------------------------------------------------------------------
// Orchestration helpers
// ---------------------------------------------------------------------------
static long long alm_count(const Handle* h) { return (long long)h->nsim * (h->lmax + 1) * (h->lmax + 1); }

static int ensure_dev_alm(Handle* h, int k) {  // sized for the capacity
  if (!h->ws_alm[k])
    HIPCHK(h, hipMalloc((void**)&h->ws_alm[k], (size_t)h->nsim_cap * (h->lmax + 1) * (h->lmax + 1) * sizeof(double)));
  return FSHT_OK;
}

// ---------------------------------------------------------------------------
// Host transfers.  Host arrays move in chunks of whole maps through two pinned
// staging buffers: host threads copy chunk k into one buffer while the DMA of
// chunk k-1 and the ring FFTs of the chunks already on the device run.
// Pinned host arrays (hipHostMalloc, e.g. torch pin_memory) skip the staging
// copy and are transferred chunk by chunk directly.
// ---------------------------------------------------------------------------
constexpr size_t kStageBytes = 128u << 20;

static bool host_pinned(const void* p) {
  hipPointerAttribute_t a;
  if (hipPointerGetAttributes(&a, p) != hipSuccess) {
    (void)hipGetLastError();  // pageable memory is reported as an error by some runtimes
    return false;
  }
  return a.type == hipMemoryTypeHost;
}

static void par_memcpy(void* dst, const void* src, size_t bytes) {
  const size_t part = 16u << 20;
  const int nt = (int)std::min<size_t>(8, std::max<size_t>(1, bytes / part));
  if (nt == 1) {
    std::memcpy(dst, src, bytes);
    return;
  }
  std::vector<std::thread> th;
  const size_t per = (bytes / nt + 63) & ~size_t(63);
  for (int t = 0; t < nt; ++t) {
    const size_t a = std::min(bytes, (size_t)t * per), b = std::min(bytes, a + per);
    if (b > a) th.emplace_back([=] { std::memcpy((char*)dst + a, (const char*)src + a, b - a); });
  }
  for (auto& x : th) x.join();
}

static int ensure_stage(Handle* h, size_t bytes) {
  bytes = std::max(bytes, kStageBytes);
  if (h->stage[0] && h->stage_bytes >= bytes) return FSHT_OK;
  HIPCHK(h, hipStreamSynchronize(h->stream));
  for (int b = 0; b < 2; ++b) {
    if (h->stage[b]) HIPCHK(h, hipHostFree(h->stage[b]));
    h->stage[b] = nullptr;
    HIPCHK(h, hipHostMalloc((void**)&h->stage[b], bytes, hipHostMallocDefault));
    if (!h->stage_ev[b]) HIPCHK(h, hipEventCreateWithFlags(&h->stage_ev[b], hipEventDisableTiming));
  }
  h->stage_bytes = bytes;
  return FSHT_OK;
}

// host -> device of n doubles in chunks of `chunk`; after(off, len) enqueues
// work on each chunk once its copy is enqueued.
template <class F>
static int h2d_chunked(Handle* h, double* dev, const double* host, size_t n, size_t chunk, F&& after) {
  const bool pinned = host_pinned(host);
  if (!pinned) {
    int rc = ensure_stage(h, chunk * sizeof(double));
    if (rc) return rc;
  }
  for (size_t off = 0, k = 0; off < n; off += chunk, ++k) {
    const size_t len = std::min(chunk, n - off);
    if (pinned) {
      HIPCHK(h, hipMemcpyAsync(dev + off, host + off, len * sizeof(double), hipMemcpyHostToDevice, h->stream));
    } else {
      const int b = (int)(k & 1);
      HIPCHK(h, hipEventSynchronize(h->stage_ev[b]));  // the buffer's previous DMA is done
      par_memcpy(h->stage[b], host + off, len * sizeof(double));
      HIPCHK(h, hipMemcpyAsync(dev + off, h->stage[b], len * sizeof(double), hipMemcpyHostToDevice, h->stream));
      HIPCHK(h, hipEventRecord(h->stage_ev[b], h->stream));
    }
    int rc = after(off, len);
    if (rc) return rc;
  }
  return FSHT_OK;
}

// device -> host of n doubles in chunks; before(off, len) enqueues the work
// that produces each chunk.  Returns when the host array holds the data.
template <class F>
static int d2h_chunked(Handle* h, double* host, const double* dev, size_t n, size_t chunk, F&& before) {
  const bool pinned = host_pinned(host);
  if (!pinned) {
    int rc = ensure_stage(h, chunk * sizeof(double));
    if (rc) return rc;
  }
  size_t poff = 0, plen = 0;
  for (size_t off = 0, k = 0; off < n; off += chunk, ++k) {
    const size_t len = std::min(chunk, n - off);
    int rc = before(off, len);
    if (rc) return rc;
    if (pinned) {
      HIPCHK(h, hipMemcpyAsync(host + off, dev + off, len * sizeof(double), hipMemcpyDeviceToHost, h->stream));
      continue;
    }
    const int b = (int)(k & 1);
    // buffer b was last used by chunk k-2, copied out in the previous iteration
    HIPCHK(h, hipMemcpyAsync(h->stage[b], dev + off, len * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipEventRecord(h->stage_ev[b], h->stream));
    if (k > 0) {
      HIPCHK(h, hipEventSynchronize(h->stage_ev[b ^ 1]));
      par_memcpy(host + poff, h->stage[b ^ 1], plen * sizeof(double));
    }
    poff = off;
    plen = len;
  }
  if (pinned) {
    HIPCHK(h, hipStreamSynchronize(h->stream));
  } else if (plen) {
    const int b = (int)(((n + chunk - 1) / chunk - 1) & 1);
    HIPCHK(h, hipEventSynchronize(h->stage_ev[b]));
    par_memcpy(host + poff, h->stage[b], plen * sizeof(double));
  }
  return FSHT_OK;
}

static size_t map_chunk(const Handle* h) {
  const size_t per = std::max<size_t>(1, kStageBytes / ((size_t)h->npix * sizeof(double)));
  return per * (size_t)h->npix;
}

static int h2d_alm(Handle* h, double* dev, const double* host) {
  return h2d_chunked(h, dev, host, (size_t)alm_count(h), kStageBytes / sizeof(double),
                     [](size_t, size_t) { return FSHT_OK; });
}
static int d2h_alm(Handle* h, double* host, const double* dev) {
  return d2h_chunked(h, host, dev, (size_t)alm_count(h), kStageBytes / sizeof(double),
                     [](size_t, size_t) { return FSHT_OK; });
}

// forward ring FFT of ncomp map sets into ws_spec (map-major).  Host maps
// are uploaded chunk by chunk, each chunk's FFTs starting once it is on the GPU.
static int forward_spectra(Handle* h, const double* m0, const double* m1, int loc) {
  const long long cnt = (long long)h->nsim * h->npix;
  if (loc) {
    const double* in[2] = {m0, m1};
    HIPCHK(h, timed(h, fsht::KC_FFT_FWD, [&] { return fsht::launch_ring_fft(h, in, h->ws_spec, 0, nullptr); }));
  } else {
    const double* in[2] = {h->ws_spec, h->ws_spec + cnt};
    const double* host[2] = {m0, m1};
    for (int c = 0; c < (h->spin ? 2 : 1); ++c) {
      int rc = h2d_chunked(h, h->ws_spec + c * cnt, host[c], (size_t)cnt, map_chunk(h), [&](size_t off, size_t len) {
        const int map0 = c * h->nsim + (int)(off / h->npix), nm = (int)(len / h->npix);
        HIPCHK(h, timed(h, fsht::KC_FFT_FWD,
                        [&] { return fsht::launch_ring_fft(h, in, h->ws_spec, 0, nullptr, map0, nm); }));
        return FSHT_OK;
      });
      if (rc) return rc;
    }
  }
  return FSHT_OK;
}

// spectra in ws_spec (map-major) -> maps.  Host maps are downloaded chunk by
// chunk behind the inverse FFTs of the following chunks.
static int inverse_spectra(Handle* h, double* m0, double* m1, int loc) {
  const long long cnt = (long long)h->nsim * h->npix;
  if (loc) {
    double* out[2] = {m0, m1};
    HIPCHK(h, timed(h, fsht::KC_FFT_INV, [&] { return fsht::launch_ring_fft(h, nullptr, h->ws_spec, 1, out); }));
    return FSHT_OK;
  }
  double* out[2] = {h->ws_spec, h->ws_spec + cnt};
  double* host[2] = {m0, m1};
  for (int c = 0; c < (h->spin ? 2 : 1); ++c) {
    int rc = d2h_chunked(h, host[c], h->ws_spec + c * cnt, (size_t)cnt, map_chunk(h), [&](size_t off, size_t len) {
      const int map0 = c * h->nsim + (int)(off / h->npix), nm = (int)(len / h->npix);
      HIPCHK(h, timed(h, fsht::KC_FFT_INV,
                      [&] { return fsht::launch_ring_fft(h, nullptr, h->ws_spec, 1, out, map0, nm); }));
      return FSHT_OK;
    });
    if (rc) return rc;
  }
  return FSHT_OK;
}

static int check_common(fsht_handle* H, int loc) {
  if (!H) return fail(nullptr, FSHT_ERR_ARG, "NULL handle");
  if (loc != FSHT_HOST && loc != FSHT_DEVICE) return fail(&H->h, FSHT_ERR_ARG, "loc must be FSHT_HOST or FSHT_DEVICE");
  H->h.err.clear();
  return FSHT_OK;
}

// analysis with optional refinement; alm outputs are device pointers.
// method 0 (Jacobi, t2alm_iter_old): every order's residual has all synthesised
// orders subtracted.  method 1 (immediate update, t2alm_iter): order m's
// residual has only the aliases m' <= m of its slot subtracted.
// Mixed precision (temperature): the same passes with int8-slice Legendre
// contractions.  Slices per pass follow how strongly the later Jacobi
// iterations damp that pass's error: the last analysis and the last two
// residual syntheses carry more.  An unrefined analysis (niter = 0) carries
// the full slice count.
// The mixed-precision state a transform of the current spin needs (built on
// first use: the spin-0 P tables of a spin-2 capable handle, the spin-2
// F1/F2 exponents and work lists).
static int ensure_mixed(Handle* h) {
  const hipError_t e = fsht::mixed_ensure(h, h->spin);
  if (e == hipErrorOutOfMemory)
    return fail(h, FSHT_ERR_NOMEM, "mixed precision: the int8 Legendre tables and sliced operands do not fit");
  if (e != hipSuccess) return fail(h, FSHT_ERR_HIP, std::string("mixed precision set-up: ") + hipGetErrorString(e));
  return FSHT_OK;
}

// Mixed precision applies to this transform: the handle runs mixed, and an
// allocation-free call (fsht_fix_eb_noalloc) only uses mixed-precision state
// that is already set up (the rest runs in FP64, as the call allocates nothing).
static bool use_mixed(const Handle* h) {
  return h->precision == FSHT_PREC_MIXED && (!h->noalloc || (h->spin ? h->mix2_ready : h->mix0_ready));
}

// One mixed-precision Legendre pass (synthesis = 0: alm <- (beta alm +) the
// analysis of the operand in FQ; 1: the m-major buffer G <- synthesis of the
// operand in AQ), range of orders by range; spin-2 ranges generate their
// F1/F2 tables first.
static int mixed_pass(Handle* h, int synthesis, double* a0, double* a1, double beta, int K, int exps) {
  if (!synthesis && exps) HIPCHK(h, fsht::launch_aqe_reset(h));
  const int nr = fsht::mixed_nrange(h);
  for (int c = 0; c < nr; ++c) {
    if (h->spin) HIPCHK(h, timed(h, fsht::KC_TABLES, [&] { return fsht::launch_ftab(h, c, synthesis, K); }));
    if (synthesis)
      HIPCHK(h, timed(h, fsht::KC_SYNTHESIS, [&] { return fsht::launch_synthesis_q(h, c, h->ws_G, K); }));
    else
      HIPCHK(h, timed(h, fsht::KC_ANALYSIS, [&] { return fsht::launch_analysis_q(h, c, a0, a1, beta, K, exps); }));
  }
  return FSHT_OK;
}

static int analysis_iter_mixed(Handle* h, double* a0, double* a1, int niter, int method) {
  int rc = ensure_mixed(h);
  if (rc) return rc;
  // error of analysis k is damped by the niter - k later iterations, that of
  // synthesis k (the residual of iteration k) by niter - k - 1
  const int sp = h->spin ? 1 : 0;
  auto lad = [&](int kind, int damping) { return h->mix_lad[sp][kind][damping < 2 ? damping : 2]; };
  HIPCHK(h, fsht::launch_fexp_reset(h));
  HIPCHK(h, timed(h, fsht::KC_GATHER, [&] { return fsht::launch_gather(h, h->ws_spec, h->ws_F0, h->d_FQe); }));
  // a single unrefined analysis carries the full slice count
  int ka = niter == 0 ? h->mix_out[sp] : lad(0, niter);
  HIPCHK(h, timed(h, fsht::KC_SLICE, [&] { return fsht::launch_quant_F(h, h->ws_F0, ka); }));
  // each analysis but the last also records alm's column exponents for the
  // next residual synthesis
  if ((rc = mixed_pass(h, 0, a0, a1, 0.0, ka, niter > 0))) return rc;
  if (niter > 0 && !h->ws_F)
    HIPCHK(h, hipMalloc((void**)&h->ws_F, (size_t)(h->G_size / (h->pol ? 1 : 2)) * sizeof(double)));
  for (int it = 0; it < niter; ++it) {
    const int ks = lad(1, niter - it - 1);
    HIPCHK(h, timed(h, fsht::KC_SLICE, [&] { return fsht::launch_quant_A(h, a0, a1, ks, 1); }));
    if ((rc = mixed_pass(h, 1, nullptr, nullptr, 0.0, ks, 0))) return rc;
    HIPCHK(h, fsht::launch_fexp_reset(h));
    HIPCHK(h, timed(h, fsht::KC_FOLD, [&] {
      return fsht::launch_gather_prefix(h, h->ws_G, h->ws_F0, h->ws_F, method == 0 ? 1 : 0, h->d_FQe);
    }));
    ka = lad(2, niter - it - 1);
    HIPCHK(h, timed(h, fsht::KC_SLICE, [&] { return fsht::launch_quant_F(h, h->ws_F, ka); }));
    if ((rc = mixed_pass(h, 0, a0, a1, 1.0, ka, it + 1 < niter))) return rc;
  }
  return FSHT_OK;
}

static int analysis_iter(Handle* h, double* a0, double* a1, int niter, int want0, int want1, int method = 0) {
  // mixed precision: temperature, and spin 2 when both E and B are wanted
  // (the F2 contractions add to what the F1 ones wrote)
  if (use_mixed(h) && (h->spin == 0 || (want0 && want1)))
    return analysis_iter_mixed(h, a0, h->spin ? a1 : nullptr, niter, method);
  HIPCHK(h, timed(h, fsht::KC_GATHER, [&] { return fsht::launch_gather(h, h->ws_spec, h->ws_F0); }));
  HIPCHK(h, timed(h, fsht::KC_ANALYSIS, [&] { return fsht::launch_analysis(h, h->ws_F0, a0, a1, 0.0, want0, want1); }));
  if (niter > 0 && !h->ws_F)
    HIPCHK(h, hipMalloc((void**)&h->ws_F, (size_t)(h->G_size / (h->pol ? 1 : 2)) * sizeof(double)));
  for (int it = 0; it < niter; ++it) {
    HIPCHK(h, timed(h, fsht::KC_SYNTHESIS, [&] { return fsht::launch_synthesis(h, a0, a1, h->ws_G); }));
    // residual + gather in one pass (the residual spectrum is never stored)
    HIPCHK(h, timed(h, fsht::KC_FOLD,
                    [&] { return fsht::launch_gather_prefix(h, h->ws_G, h->ws_F0, h->ws_F, method == 0 ? 1 : 0); }));
    HIPCHK(h, timed(h, fsht::KC_ANALYSIS, [&] { return fsht::launch_analysis(h, h->ws_F, a0, a1, 1.0, want0, want1); }));
  }
  return FSHT_OK;
}

extern "C" {

static int t2alm_impl(fsht_handle* H, const double* maps, double* alms, int niter, int loc, int method);

int fsht_t2alm(fsht_handle* H, const double* maps, double* alms, int niter, int loc) {
  return t2alm_impl(H, maps, alms, niter, loc, 0);
}

int fsht_t2alm_iter(fsht_handle* H, const double* maps, double* alms, int niter, int loc) {
  return t2alm_impl(H, maps, alms, niter, loc, 1);
}

static int t2alm_impl(fsht_handle* H, const double* maps, double* alms, int niter, int loc, int method) {
  int rc = check_common(H, loc);
  if (rc) return rc;
  Handle* h = &H->h;
  if (!maps) return fail(h, FSHT_ERR_ARG, "maps is NULL");
  if (niter < 0) return fail(h, FSHT_ERR_ARG, "niter must be >= 0");
  h->spin = 0;
  if ((rc = forward_spectra(h, maps, nullptr, loc))) return rc;
  double* dst = alms;
  const bool keep = (alms == nullptr);
  if (keep || !loc) {
    if ((rc = ensure_dev_alm(h, 0))) return rc;
    dst = h->ws_alm[0];
  }
  if ((rc = analysis_iter(h, dst, nullptr, niter, 1, 0, method))) return rc;
  if (!keep && !loc) {
    if ((rc = d2h_alm(h, alms, dst))) return rc;
  }
  h->have_dev_alm = keep ? 1 : 0;
  return FSHT_OK;
}

int fsht_get_alm(fsht_handle* H, double* alms, int loc) {
  int rc = check_common(H, loc);
  if (rc) return rc;
  Handle* h = &H->h;
  if (!h->have_dev_alm || !h->ws_alm[0]) return fail(h, FSHT_ERR_STATE, "no device-resident alm: call t2alm with alms=NULL first");
  if (loc) {
    HIPCHK(h, hipMemcpyAsync(alms, h->ws_alm[0], alm_count(h) * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
  } else if ((rc = d2h_alm(h, alms, h->ws_alm[0]))) {
    return rc;
  }
  return FSHT_OK;
}

int fsht_alm2t(fsht_handle* H, const double* alms, double* maps, int loc) {
  int rc = check_common(H, loc);
  if (rc) return rc;
  Handle* h = &H->h;
  if (!alms || !maps) return fail(h, FSHT_ERR_ARG, "alms/maps is NULL");
  h->spin = 0;
  const double* a = alms;
  if (!loc) {
    if ((rc = ensure_dev_alm(h, 0))) return rc;
    if ((rc = h2d_alm(h, h->ws_alm[0], alms))) return rc;
    a = h->ws_alm[0];
    h->have_dev_alm = 0;
  }
  if (use_mixed(h)) {
    if ((rc = ensure_mixed(h))) return rc;
    HIPCHK(h, timed(h, fsht::KC_SLICE, [&] { return fsht::launch_quant_A(h, a, nullptr, h->mix_out[0], 0); }));
    if ((rc = mixed_pass(h, 1, nullptr, nullptr, 0.0, h->mix_out[0], 0))) return rc;
  } else {
    HIPCHK(h, timed(h, fsht::KC_SYNTHESIS, [&] { return fsht::launch_synthesis(h, a, nullptr, h->ws_G); }));
  }
  HIPCHK(h, timed(h, fsht::KC_FOLD, [&] { return fsht::launch_fold(h, h->ws_G, h->ws_spec); }));
  if ((rc = inverse_spectra(h, maps, nullptr, loc))) return rc;
  if (!loc) HIPCHK(h, hipStreamSynchronize(h->stream));
  return FSHT_OK;
}

static int qu_forward(fsht_handle* H, const double* Q, const double* U, double* E, double* B, int niter, int loc,
                      int want0, int want1) {
  int rc = check_common(H, loc);
  if (rc) return rc;
  Handle* h = &H->h;
  if (!h->pol) return fail(h, FSHT_ERR_STATE, "spin-2 transforms need a handle created with pol=1");
  if (!Q || !U) return fail(h, FSHT_ERR_ARG, "Q/U is NULL");
  if ((want0 && !E) || (want1 && !B)) return fail(h, FSHT_ERR_ARG, "output alm is NULL");
  if (niter < 0) return fail(h, FSHT_ERR_ARG, "niter must be >= 0");
  h->spin = 2;
  if ((rc = forward_spectra(h, Q, U, loc))) return rc;
  double* e = E;
  double* b = B;
  if (!loc) {
    if ((rc = ensure_dev_alm(h, 0))) return rc;
    if ((rc = ensure_dev_alm(h, 1))) return rc;
    e = h->ws_alm[0];
    b = h->ws_alm[1];
    h->have_dev_alm = 0;
  }
  if (!want0) e = b;  // never written
  if (!want1) b = e;
  if ((rc = analysis_iter(h, e, b, niter, want0, want1))) return rc;
  if (!loc) {
    if (want0 && (rc = d2h_alm(h, E, e))) return rc;
    if (want1 && (rc = d2h_alm(h, B, b))) return rc;
  }
  return FSHT_OK;
}

int fsht_qu2eb(fsht_handle* H, const double* Q, const double* U, double* E, double* B, int niter, int loc) {
  return qu_forward(H, Q, U, E, B, niter, loc, 1, 1);
}
int fsht_qu2e(fsht_handle* H, const double* Q, const double* U, double* E, int loc) {
  return qu_forward(H, Q, U, E, nullptr, 0, loc, 1, 0);
}
int fsht_qu2b(fsht_handle* H, const double* Q, const double* U, double* B, int loc) {
  return qu_forward(H, Q, U, nullptr, B, 0, loc, 0, 1);
}

int fsht_eb2qu(fsht_handle* H, const double* E, const double* B, double* Q, double* U, int loc) {
  int rc = check_common(H, loc);
  if (rc) return rc;
  Handle* h = &H->h;
  if (!h->pol) return fail(h, FSHT_ERR_STATE, "spin-2 transforms need a handle created with pol=1");
  if (!E || !B || !Q || !U) return fail(h, FSHT_ERR_ARG, "NULL argument");
  h->spin = 2;
  const double* e = E;
  const double* b = B;
  if (!loc) {
    if ((rc = ensure_dev_alm(h, 0))) return rc;
    if ((rc = ensure_dev_alm(h, 1))) return rc;
    if ((rc = h2d_alm(h, h->ws_alm[0], E))) return rc;
    if ((rc = h2d_alm(h, h->ws_alm[1], B))) return rc;
    e = h->ws_alm[0];
    b = h->ws_alm[1];
    h->have_dev_alm = 0;
  }
  if (use_mixed(h)) {
    if ((rc = ensure_mixed(h))) return rc;
    HIPCHK(h, timed(h, fsht::KC_SLICE, [&] { return fsht::launch_quant_A(h, e, b, h->mix_out[1], 0); }));
    if ((rc = mixed_pass(h, 1, nullptr, nullptr, 0.0, h->mix_out[1], 0))) return rc;
  } else {
    HIPCHK(h, timed(h, fsht::KC_SYNTHESIS, [&] { return fsht::launch_synthesis(h, e, b, h->ws_G); }));
  }
  HIPCHK(h, timed(h, fsht::KC_FOLD, [&] { return fsht::launch_fold(h, h->ws_G, h->ws_spec); }));
  if ((rc = inverse_spectra(h, Q, U, loc))) return rc;
  if (!loc) HIPCHK(h, hipStreamSynchronize(h->stream));
  return FSHT_OK;
}

// ---------------------------------------------------------------------------
// EB-leakage cleaning on a masked sky (src/sht_main.f90:801-941):
//   Q, U *= mask (in place, as the reference does)
//   (E, B)  = qu2eb(Q, U, niter)
//   (Q', U') = E-only synthesis of E, masked again
//   (., B') = qu2eb(Q', U', niter)           -- harmonic leakage template
//   Bmap0 = alm2t(B), Btpl = alm2t(B')
//   per map, over the pixels with mask == 1: fit Bmap0 = a0 + a1*Btpl and
//   return the residual there (0 elsewhere).
// Temporaries are allocated for the call and released at its end.
// ---------------------------------------------------------------------------
namespace {
struct DevBuf {
  double* p = nullptr;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
};
}  // namespace

// The fix_eb sequence on device arrays with given temporaries: alm sets dE,
// dB2, dZ and map sets dQ2, dU2 (ab, bt, cf: almB, the template, the fit).
static int fix_eb_device(fsht_handle* H, double* q, double* u, const double* mk, double* bm, double* ab, double* bt,
                         double* cf, int niter, double* dE, double* dB2, double* dZ, double* dQ2, double* dU2) {
  Handle* h = &H->h;
  int rc;
  HIPCHK(h, fsht::launch_zero(dZ, alm_count(h), h->stream));
  HIPCHK(h, fsht::launch_mask_maps(q, mk, h->npix, h->nsim, h->stream));
  HIPCHK(h, fsht::launch_mask_maps(u, mk, h->npix, h->nsim, h->stream));
  if ((rc = fsht_qu2eb(H, q, u, dE, ab, niter, FSHT_DEVICE))) return rc;
  if ((rc = fsht_eb2qu(H, dE, dZ, dQ2, dU2, FSHT_DEVICE))) return rc;
  HIPCHK(h, fsht::launch_mask_maps(dQ2, mk, h->npix, h->nsim, h->stream));
  HIPCHK(h, fsht::launch_mask_maps(dU2, mk, h->npix, h->nsim, h->stream));
  if ((rc = fsht_qu2eb(H, dQ2, dU2, dE, dB2, niter, FSHT_DEVICE))) return rc;
  if ((rc = fsht_alm2t(H, ab, dQ2, FSHT_DEVICE))) return rc;  // B map with leakage
  if ((rc = fsht_alm2t(H, dB2, bt, FSHT_DEVICE))) return rc;  // leakage template
  HIPCHK(h, fsht::launch_linfit_resi(bt, dQ2, mk, bm, cf, h->npix, h->nsim, h->stream));
  return FSHT_OK;
}

int fsht_fix_eb(fsht_handle* H, double* Q, double* U, const double* mask, double* Bmap, double* almB, double* Btpl,
                double* coef, int niter, int loc) {
  int rc = check_common(H, loc);
  if (rc) return rc;
  Handle* h = &H->h;
  if (!h->pol) return fail(h, FSHT_ERR_STATE, "fix_eb needs a handle created with pol=1");
  if (!Q || !U || !mask || !Bmap) return fail(h, FSHT_ERR_ARG, "NULL argument");
  if (niter < 0) return fail(h, FSHT_ERR_ARG, "niter must be >= 0");
  const long long nm = (long long)h->nsim * h->npix;
  const size_t mbytes = (size_t)nm * sizeof(double), abytes = (size_t)alm_count(h) * sizeof(double);
  // temporaries outlive the work: the stream is drained before they are freed,
  // on the error paths too
  DevBuf dQ, dU, dmask, dB, dE, dB2, dZ, dQ2, dU2, dBmap, dBtpl, dcoef;
  rc = [&]() -> int {
    auto alloc = [&](DevBuf& b, size_t bytes) -> hipError_t { return hipMalloc((void**)&b.p, std::max<size_t>(bytes, 8)); };
    double *q = Q, *u = U, *bm = Bmap, *bt = Btpl, *ab = almB, *cf = coef;
    const double* mk = mask;
    if (!loc) {
      HIPCHK(h, alloc(dQ, mbytes));
      HIPCHK(h, alloc(dU, mbytes));
      HIPCHK(h, alloc(dmask, (size_t)h->npix * sizeof(double)));
      HIPCHK(h, alloc(dBmap, mbytes));
      HIPCHK(h, hipMemcpyAsync(dQ.p, Q, mbytes, hipMemcpyHostToDevice, h->stream));
      HIPCHK(h, hipMemcpyAsync(dU.p, U, mbytes, hipMemcpyHostToDevice, h->stream));
      HIPCHK(h, hipMemcpyAsync(dmask.p, mask, (size_t)h->npix * sizeof(double), hipMemcpyHostToDevice, h->stream));
      q = dQ.p;
      u = dU.p;
      mk = dmask.p;
      bm = dBmap.p;
      bt = nullptr;
      ab = nullptr;
      cf = nullptr;
    }
    if (!bt) {
      HIPCHK(h, alloc(dBtpl, mbytes));
      bt = dBtpl.p;
    }
    if (!ab) {
      HIPCHK(h, alloc(dB, abytes));
      ab = dB.p;
    }
    if (!cf) {
      HIPCHK(h, alloc(dcoef, 2 * (size_t)h->nsim * sizeof(double)));
      cf = dcoef.p;
    }
    HIPCHK(h, alloc(dE, abytes));
    HIPCHK(h, alloc(dB2, abytes));
    HIPCHK(h, alloc(dZ, abytes));
    HIPCHK(h, alloc(dQ2, mbytes));
    HIPCHK(h, alloc(dU2, mbytes));
    int rc2 = fix_eb_device(H, q, u, mk, bm, ab, bt, cf, niter, dE.p, dB2.p, dZ.p, dQ2.p, dU2.p);
    if (rc2) return rc2;
    if (!loc) {
      HIPCHK(h, hipMemcpyAsync(Q, dQ.p, mbytes, hipMemcpyDeviceToHost, h->stream));
      HIPCHK(h, hipMemcpyAsync(U, dU.p, mbytes, hipMemcpyDeviceToHost, h->stream));
      HIPCHK(h, hipMemcpyAsync(Bmap, bm, mbytes, hipMemcpyDeviceToHost, h->stream));
      if (Btpl) HIPCHK(h, hipMemcpyAsync(Btpl, bt, mbytes, hipMemcpyDeviceToHost, h->stream));
      if (almB) HIPCHK(h, hipMemcpyAsync(almB, ab, abytes, hipMemcpyDeviceToHost, h->stream));
      if (coef) HIPCHK(h, hipMemcpyAsync(coef, cf, 2 * (size_t)h->nsim * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    }
    return FSHT_OK;
  }();
  const hipError_t se = hipStreamSynchronize(h->stream);
  if (rc == FSHT_OK && se != hipSuccess) return fail(h, FSHT_ERR_HIP, std::string("fix_eb: ") + hipGetErrorString(se));
  return rc;
}

int fsht_fix_eb_noalloc(fsht_handle* H, double* Q, double* U, const double* mask, double* Bmap, double* almB,
                        double* Btpl, double* coef, int niter) {
  int rc = check_common(H, FSHT_DEVICE);
  if (rc) return rc;
  Handle* h = &H->h;
  if (!h->pol) return fail(h, FSHT_ERR_STATE, "fix_eb needs a handle created with pol=1");
  if (!Q || !U || !mask || !Bmap) return fail(h, FSHT_ERR_ARG, "NULL argument");
  if (niter < 0) return fail(h, FSHT_ERR_ARG, "niter must be >= 0");
  if (niter > 0 && !h->ws_F)
    return fail(h, FSHT_ERR_STATE, "niter > 0 needs a handle created with niter > 0 (its workspace is set up then)");
  if (!h->scratch) return fail(h, FSHT_ERR_STATE, "no scratch address for this device");
  // carve the temporaries out of the scratch, 256-byte aligned
  const size_t abytes = (size_t)alm_count(h) * sizeof(double);
  const size_t mbytes = (size_t)h->nsim * (size_t)h->npix * sizeof(double);
  auto up = [](size_t b) { return (b + 255) & ~size_t(255); };
  const size_t need = 3 * up(abytes) + 2 * up(mbytes) + (almB ? 0 : up(abytes)) + (Btpl ? 0 : up(mbytes)) +
                      (coef ? 0 : up(2 * (size_t)h->nsim * sizeof(double)));
  if (need > kScratchBytes)
    return fail(h, FSHT_ERR_ARG,
                "fix_eb_noalloc: this configuration needs " + std::to_string(need) + " bytes of scratch, more than the " +
                    std::to_string(kScratchBytes) + " available");
  char* p = static_cast<char*>(h->scratch);
  auto take = [&](size_t b) {
    double* r = reinterpret_cast<double*>(p);
    p += up(b);
    return r;
  };
  double* dE = take(abytes);
  double* dB2 = take(abytes);
  double* dZ = take(abytes);
  double* dQ2 = take(mbytes);
  double* dU2 = take(mbytes);
  double* ab = almB ? almB : take(abytes);
  double* bt = Btpl ? Btpl : take(mbytes);
  double* cf = coef ? coef : take(2 * (size_t)h->nsim * sizeof(double));
  // no profiling events either: recording them would create events
  const int prof = h->profiling;
  h->profiling = 0;
  h->noalloc = 1;
  rc = fix_eb_device(H, Q, U, mask, Bmap, ab, bt, cf, niter, dE, dB2, dZ, dQ2, dU2);
  h->noalloc = 0;
  h->profiling = prof;
  return rc;
}

int fsht_host_alloc(unsigned long long bytes, void** out) {
  if (!out) return fail(nullptr, FSHT_ERR_ARG, "out is NULL");
  *out = nullptr;
  hipError_t e = hipHostMalloc(out, std::max<unsigned long long>(bytes, 8), hipHostMallocDefault);
  if (e != hipSuccess) return fail(nullptr, FSHT_ERR_NOMEM, std::string("hipHostMalloc: ") + hipGetErrorString(e));
  return FSHT_OK;
}

int fsht_host_free(void* p) {
  if (!p) return FSHT_OK;
  hipError_t e = hipHostFree(p);
  return e == hipSuccess ? FSHT_OK : fail(nullptr, FSHT_ERR_HIP, hipGetErrorString(e));
}

int fsht_profile(fsht_handle* H, int enable) {
  if (!H) return fail(nullptr, FSHT_ERR_ARG, "NULL handle");
  Handle* h = &H->h;
  HIPCHK(h, hipStreamSynchronize(h->stream));
  for (auto& t : h->timed) {
    h->event_pool.push_back(t.a);
    h->event_pool.push_back(t.b);
  }
  h->timed.clear();
  h->profiling = enable ? 1 : 0;
  return FSHT_OK;
}

int fsht_profile_read(fsht_handle* H, double* out, int n) {
  if (!H || !out) return fail(nullptr, FSHT_ERR_ARG, "NULL argument");
  Handle* h = &H->h;
  HIPCHK(h, hipStreamSynchronize(h->stream));
  double acc[2 * fsht::KC_COUNT] = {0};
  for (auto& t : h->timed) {
    float ms = 0.f;
    HIPCHK(h, hipEventElapsedTime(&ms, t.a, t.b));
    acc[2 * t.cls] += 1.0;
    acc[2 * t.cls + 1] += ms;
  }
  for (int i = 0; i < n && i < 2 * fsht::KC_COUNT; ++i) out[i] = acc[i];
  return FSHT_OK;
}

// ---- alm utilities (device pointers, caller's stream; NULL = default stream)
int fsht_norm_alm(const double* alm_in, double* alm_out, const double* fac, const double* shift, double m0, int lmax,
                  int nsim, void* stream) {
  if (!alm_in || !alm_out || !fac || !shift || lmax < 0 || nsim < 1) return fail(nullptr, FSHT_ERR_ARG, "bad argument");
  hipError_t e = fsht::launch_norm_alm(alm_in, alm_out, fac, shift, m0, lmax, nsim, (hipStream_t)stream);
  return e == hipSuccess ? FSHT_OK : fail(nullptr, FSHT_ERR_HIP, hipGetErrorString(e));
}

int fsht_filt_alm(double* alm, const double* win, int lmax, int nsim, void* stream) {
  if (!alm || !win || lmax < 0 || nsim < 1) return fail(nullptr, FSHT_ERR_ARG, "bad argument");
  hipError_t e = fsht::launch_filt_alm(alm, win, lmax, nsim, (hipStream_t)stream);
  return e == hipSuccess ? FSHT_OK : fail(nullptr, FSHT_ERR_HIP, hipGetErrorString(e));
}

// Square packing <-> healpy order.  Host arrays (loc = FSHT_HOST) are staged
// through device buffers allocated for the call; the call then returns with
// the result in `dst`.
static int convert_healpy(const double* src, double* dst, int lmax, int nsim, int inverse, int loc, void* stream) {
  if (!src || !dst || lmax < 0 || nsim < 1 || (loc != FSHT_HOST && loc != FSHT_DEVICE))
    return fail(nullptr, FSHT_ERR_ARG, "bad argument");
  const size_t L = (size_t)lmax + 1;
  const size_t n_sq = L * L * (size_t)nsim, n_hp = L * (L + 1) * (size_t)nsim;  // doubles
  const size_t n_src = inverse ? n_hp : n_sq, n_dst = inverse ? n_sq : n_hp;
  hipStream_t st = (hipStream_t)stream;
  hipError_t e;
  if (loc == FSHT_DEVICE) {
    e = fsht::launch_convert_healpy(src, dst, lmax, nsim, inverse, st);
  } else {
    DevBuf ds, dd;
    e = hipMalloc((void**)&ds.p, n_src * sizeof(double));
    if (e == hipSuccess) e = hipMalloc((void**)&dd.p, n_dst * sizeof(double));
    if (e == hipSuccess) e = hipMemcpyAsync(ds.p, src, n_src * sizeof(double), hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = fsht::launch_convert_healpy(ds.p, dd.p, lmax, nsim, inverse, st);
    if (e == hipSuccess) e = hipMemcpyAsync(dst, dd.p, n_dst * sizeof(double), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
  }
  return e == hipSuccess ? FSHT_OK : fail(nullptr, FSHT_ERR_HIP, hipGetErrorString(e));
}

int fsht_convert_alm_healpy(const double* alm, double* alm_hp, int lmax, int nsim, int loc, void* stream) {
  return convert_healpy(alm, alm_hp, lmax, nsim, 0, loc, stream);
}

int fsht_convert_alm_healpy_inv(const double* alm_hp, double* alm, int lmax, int nsim, int loc, void* stream) {
  return convert_healpy(alm_hp, alm, lmax, nsim, 1, loc, stream);
}

int fsht_alm2cl(const double* almT, const double* almE, const double* almB, double* cl, int lmin, int lmax, int nsim,
                void* stream) {
  if (!almT || !cl || lmax < 0 || nsim < 1 || lmin < 0) return fail(nullptr, FSHT_ERR_ARG, "bad argument");
  if ((almE == nullptr) != (almB == nullptr)) return fail(nullptr, FSHT_ERR_ARG, "almE and almB go together");
  hipError_t e = fsht::launch_alm2cl(almT, almE, almB, cl, lmin, lmax, nsim, almE ? 6 : 1, (hipStream_t)stream);
  return e == hipSuccess ? FSHT_OK : fail(nullptr, FSHT_ERR_HIP, hipGetErrorString(e));
}

int fsht_cl2alm(const double* cl, double* almT, double* almE, double* almB, int lmin, int lmax, int nsim,
                unsigned long long seed, void* stream) {
  if (!cl || !almT || lmax < 0 || nsim < 1 || lmin < 0) return fail(nullptr, FSHT_ERR_ARG, "bad argument");
  if ((almE == nullptr) != (almB == nullptr)) return fail(nullptr, FSHT_ERR_ARG, "almE and almB go together");
  hipError_t e = fsht::launch_cl2alm(cl, almT, almE, almB, lmin, lmax, nsim, almE ? 3 : 1, seed, (hipStream_t)stream);
  return e == hipSuccess ? FSHT_OK : fail(nullptr, FSHT_ERR_HIP, hipGetErrorString(e));
}

}  // extern "C"
