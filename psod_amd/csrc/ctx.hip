// ctx.hip -- context, errors, scratch, scan and radix sort primitives.
#include <cstdarg>
#include <cstdlib>
#include <cstring>

#include "psod_internal.h"

static thread_local char g_err[1024] = "";

int psod_set_error(int code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return code;
}

extern "C" const char *psod_last_error(void) { return g_err; }

extern "C" int psod_create(int device, psod_ctx **out) {
  if (!out) return psod_set_error(PSOD_E_ARG, "psod_create: out is NULL");
  *out = nullptr;
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n == 0)
    return psod_set_error(PSOD_E_DEVICE,
                          "psod_create: no HIP device (%s)",
                          hipGetErrorString(e));
  if (device < 0 || device >= n)
    return psod_set_error(PSOD_E_ARG, "psod_create: device %d of %d", device,
                          n);
  PSOD_HIP(hipSetDevice(device));
  psod_ctx *c = (psod_ctx *)calloc(1, sizeof(psod_ctx));
  c->device = device;
  hipError_t es = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (es != hipSuccess) {
    free(c);
    return psod_set_error(PSOD_E_DEVICE, "hipStreamCreate: %s",
                          hipGetErrorString(es));
  }
  if (hipEventCreate(&c->ev[0]) != hipSuccess ||
      hipEventCreate(&c->ev[1]) != hipSuccess ||
      hipEventCreateWithFlags(&c->fork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c->join, hipEventDisableTiming) != hipSuccess ||
      hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking) != hipSuccess) {
    hipStreamDestroy(c->stream);
    free(c);
    return psod_set_error(PSOD_E_DEVICE, "hipEventCreate failed");
  }
  // the fixed tile list of psod_cic_deposit_noalloc on this device
  const int rs = psod_static_scratch_init(device);
  if (rs != PSOD_OK) {
    hipEventDestroy(c->ev[0]);
    hipEventDestroy(c->ev[1]);
    hipEventDestroy(c->fork);
    hipEventDestroy(c->join);
    hipStreamDestroy(c->side);
    hipStreamDestroy(c->stream);
    free(c);
    return rs;
  }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess) {
    c->num_cus = prop.multiProcessorCount;
    snprintf(c->device_name, sizeof(c->device_name), "%s", prop.gcnArchName);
  } else {
    c->num_cus = 256;
    snprintf(c->device_name, sizeof(c->device_name), "unknown");
  }
  // names are single words in the tuning file
  for (char *q = c->device_name; *q; q++)
    if (*q == ' ' || *q == '\t') *q = '_';
  *out = c;
  if (const char *tf = getenv("PSOD_TUNE_FILE")) {
    int32_t nl = 0;
    const int rl = psod_load_tunables(c, tf, &nl);
    if (rl != PSOD_OK) {
      psod_destroy(c);
      *out = nullptr;
      return rl;
    }
  }
  return PSOD_OK;
}

extern "C" int psod_destroy(psod_ctx *c) {
  if (!c) return PSOD_OK;
  hipSetDevice(c->device);
  hipStreamSynchronize(c->stream);
  hipStreamSynchronize(c->side);
  psod_graph_release(c);
  for (int s = 0; s < S_NSLOT; s++)
    if (c->slot[s]) hipFree(c->slot[s]);
  hipEventDestroy(c->ev[0]);
  hipEventDestroy(c->ev[1]);
  hipEventDestroy(c->fork);
  hipEventDestroy(c->join);
  hipStreamDestroy(c->side);
  hipStreamDestroy(c->stream);
  free(c);
  return PSOD_OK;
}

extern "C" int psod_synchronize(psod_ctx *c) {
  if (!c) return psod_set_error(PSOD_E_ARG, "null context");
  PSOD_HIP(hipStreamSynchronize(c->stream));
  PSOD_HIP(hipStreamSynchronize(c->side));
  return PSOD_OK;
}

extern "C" void *psod_stream(psod_ctx *c) { return c ? (void *)c->stream : nullptr; }

extern "C" int64_t psod_scratch_bytes(psod_ctx *c) {
  int64_t s = 0;
  if (c)
    for (int i = 0; i < S_NSLOT; i++) s += (int64_t)c->cap[i];
  return s;
}

extern "C" int psod_sod_stats(psod_ctx *c, int64_t *out, int32_t n) {
  if (!c || !out || n < 0) return psod_set_error(PSOD_E_ARG, "psod_sod_stats: bad arguments");
  for (int i = 0; i < n && i < 13; i++) out[i] = c->sod_stat[i];
  return PSOD_OK;
}

int psod_use_device(psod_ctx *c) {
  if (!c) return psod_set_error(PSOD_E_ARG, "null context");
  if (c->capturing) return PSOD_OK;  // the capturing call made it current
  PSOD_HIP(hipSetDevice(c->device));
  return PSOD_OK;
}

int psod_finish(psod_ctx *c) {
  if (c->capturing) return PSOD_OK;
  PSOD_HIP(hipStreamSynchronize(c->stream));
  return PSOD_OK;
}

void *psod_scratch(psod_ctx *c, int slot, size_t bytes) {
  if (bytes == 0) bytes = 256;
  if (c->cap[slot] >= bytes) return c->slot[slot];
  if (c->capturing) {
    psod_set_error(PSOD_E_CAPACITY, "scratch slot %d would grow from %zu to "
                   "%zu bytes while the pass is captured", slot, c->cap[slot],
                   bytes);
    return nullptr;
  }
  // a captured pass holds the old buffer's address: it is dropped (and
  // captured again by the next psod_pass_graph call)
  psod_graph_release(c);
  if (c->slot[slot]) {
    hipStreamSynchronize(c->stream);
    hipStreamSynchronize(c->side);
    hipFree(c->slot[slot]);
    c->slot[slot] = nullptr;
    c->cap[slot] = 0;
  }
  size_t want = bytes + bytes / 8;
  void *p = nullptr;
  if (hipMalloc(&p, want) != hipSuccess) {
    (void)hipGetLastError();
    if (hipMalloc(&p, bytes) != hipSuccess) {
      (void)hipGetLastError();
      psod_set_error(PSOD_E_NOMEM, "device allocation of %zu bytes failed",
                     bytes);
      return nullptr;
    }
    want = bytes;
  }
  c->slot[slot] = p;
  c->cap[slot] = want;
  return p;
}

// ---------------------------------------------------------------------------
// exclusive scan of int32, 256 threads x 16 elements per block
// ---------------------------------------------------------------------------
constexpr int SCAN_T = 256, SCAN_V = 16, SCAN_TILE = SCAN_T * SCAN_V;

__device__ inline int32_t block_excl_scan(int32_t v, int32_t *total) {
  __shared__ int32_t wsum[SCAN_T / 64];
  const int l = threadIdx.x & 63, w = threadIdx.x >> 6;
  int32_t inc = wave_incl_scan(v);
  if (l == 63) wsum[w] = inc;
  __syncthreads();
  int32_t off = 0, tot = 0;
  for (int i = 0; i < SCAN_T / 64; i++) {
    if (i < w) off += wsum[i];
    tot += wsum[i];
  }
  __syncthreads();
  if (total) *total = tot;
  return off + inc - v;
}

__global__ void __launch_bounds__(SCAN_T)
k_scan_reduce(const int32_t *__restrict__ in, int64_t n,
              int32_t *__restrict__ part) {
  const int64_t base = (int64_t)blockIdx.x * SCAN_TILE;
  int32_t s = 0;
  for (int k = 0; k < SCAN_V; k++) {
    const int64_t i = base + (int64_t)k * SCAN_T + threadIdx.x;
    if (i < n) s += in[i];
  }
  int32_t tot;
  (void)block_excl_scan(s, &tot);
  if (threadIdx.x == 0) part[blockIdx.x] = tot;
}

__global__ void __launch_bounds__(SCAN_T)
k_scan_apply(const int32_t *__restrict__ in, int64_t n,
             const int32_t *__restrict__ partscan, int32_t *__restrict__ out) {
  const int64_t base = (int64_t)blockIdx.x * SCAN_TILE +
                       (int64_t)threadIdx.x * SCAN_V;
  int32_t v[SCAN_V];
  int32_t s = 0;
  for (int k = 0; k < SCAN_V; k++) {
    const int64_t i = base + k;
    v[k] = (i < n) ? in[i] : 0;
    s += v[k];
  }
  int32_t ex = block_excl_scan(s, nullptr) + (partscan ? partscan[blockIdx.x] : 0);
  for (int k = 0; k < SCAN_V; k++) {
    const int64_t i = base + k;
    if (i < n) out[i] = ex;
    ex += v[k];
  }
}

__global__ void k_scan_total(const int32_t *in, const int32_t *out, int64_t n,
                             int32_t *total) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *total = out[n - 1] + in[n - 1];
}

int psod_scan_i32(psod_ctx *c, const int32_t *in, int32_t *out, int64_t n,
                  int32_t *total) {
  if (n <= 0) return PSOD_OK;
  if (in == out)
    return psod_set_error(PSOD_E_ARG, "psod_scan_i32 is out of place");
  // level sizes: n, ceil(n/T), ceil(n/T^2), ... down to one tile
  int64_t len[8];
  int nlev = 0;
  len[nlev++] = n;
  while (len[nlev - 1] > SCAN_TILE) {
    if (nlev == 8) return psod_set_error(PSOD_E_ARG, "scan too long");
    len[nlev] = (len[nlev - 1] + SCAN_TILE - 1) / SCAN_TILE;
    nlev++;
  }
  // one buffer for the partials (part) and their scans (ps) of every level
  int64_t need = 0;
  for (int l = 1; l < nlev; l++) need += 2 * (len[l] + 64);
  int32_t *buf = (int32_t *)psod_scratch(c, S_SCAN, (size_t)(need + 64) * 4);
  if (!buf) return PSOD_E_NOMEM;
  const int32_t *lin[8];
  int32_t *lout[8];
  lin[0] = in;
  lout[0] = out;
  int64_t o = 0;
  for (int l = 1; l < nlev; l++) {
    lin[l] = buf + o;
    o += len[l] + 64;
    lout[l] = buf + o;
    o += len[l] + 64;
  }
  // reduce down the levels
  for (int l = 0; l + 1 < nlev; l++) {
    k_scan_reduce<<<(unsigned)len[l + 1], SCAN_T, 0, c->stream>>>(
        lin[l], len[l], (int32_t *)lin[l + 1]);
    PSOD_LAUNCH_CHECK();
  }
  // top level in one tile, then apply back up
  k_scan_apply<<<1, SCAN_T, 0, c->stream>>>(lin[nlev - 1], len[nlev - 1],
                                             nullptr, lout[nlev - 1]);
  PSOD_LAUNCH_CHECK();
  for (int l = nlev - 2; l >= 0; l--) {
    k_scan_apply<<<(unsigned)len[l + 1], SCAN_T, 0, c->stream>>>(
        lin[l], len[l], lout[l + 1], lout[l]);
    PSOD_LAUNCH_CHECK();
  }
  if (total) {
    k_scan_total<<<1, 64, 0, c->stream>>>(in, out, n, total);
    PSOD_LAUNCH_CHECK();
  }
  return PSOD_OK;
}

// ---------------------------------------------------------------------------
// LSD radix sort, 8-bit digits, u64 keys + u32 values, stable.
// Per pass: (1) per-tile digit histograms stored digit-major, (2) exclusive
// scan over [digit][tile], (3) stable scatter.  A tile is 256 threads x 16
// keys; inside a tile each thread owns 16 consecutive keys and the ranks of
// equal digits are formed with per-wave ballots in thread order.
// ---------------------------------------------------------------------------
constexpr int RS_T = 256, RS_V = 16, RS_TILE = RS_T * RS_V, RS_R = 256;

__global__ void __launch_bounds__(RS_T)
k_rs_hist(const uint64_t *__restrict__ keys, int64_t n,
          const int32_t *__restrict__ dn, int shift,
          int32_t *__restrict__ hist, int64_t ntile) {
  __shared__ int32_t h[RS_R];
  if (dn) n = min(n, (int64_t)*dn);
  h[threadIdx.x] = 0;
  __syncthreads();
  const int64_t base = (int64_t)blockIdx.x * RS_TILE;
  for (int k = 0; k < RS_V; k++) {
    const int64_t i = base + (int64_t)k * RS_T + threadIdx.x;
    if (i < n) atomicAdd(&h[(keys[i] >> shift) & 0xFF], 1);
  }
  __syncthreads();
  hist[(int64_t)threadIdx.x * ntile + blockIdx.x] = h[threadIdx.x];
}

__global__ void __launch_bounds__(RS_T)
k_rs_scatter(const uint64_t *__restrict__ kin, const uint32_t *__restrict__ vin,
             int64_t n, const int32_t *__restrict__ dn, int shift,
             const int32_t *__restrict__ offs, int64_t ntile,
             uint64_t *__restrict__ kout, uint32_t *__restrict__ vout) {
  if (dn) n = min(n, (int64_t)*dn);
  // Process the tile in RS_V rounds of 256 consecutive keys (one per
  // thread, thread order == index order), keeping a running count per digit.
  __shared__ int32_t run[RS_R];
  __shared__ int32_t wcnt[RS_T / 64][RS_R];
  run[threadIdx.x] = offs[(int64_t)threadIdx.x * ntile + blockIdx.x];
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int64_t base = (int64_t)blockIdx.x * RS_TILE;
  for (int k = 0; k < RS_V; k++) {
    for (int d = threadIdx.x; d < RS_R * (RS_T / 64); d += RS_T)
      (&wcnt[0][0])[d] = 0;
    __syncthreads();
    const int64_t i = base + (int64_t)k * RS_T + threadIdx.x;
    const bool ok = i < n;
    uint64_t key = ok ? kin[i] : 0;
    uint32_t val = ok ? vin[i] : 0;
    const uint32_t dg = (uint32_t)((key >> shift) & 0xFF);
    // lanes of this wave with the same digit
    uint64_t same = ballot64(ok);
    for (int b = 0; b < 8; b++) {
      const uint64_t bb = ballot64(ok && ((dg >> b) & 1));
      same &= ((dg >> b) & 1) ? bb : ~bb;
    }
    const uint64_t lower = (l == 0) ? 0ull : (~0ull >> (64 - l));
    const int rank_w = __popcll(same & lower);
    const int cnt_w = __popcll(same);
    if (ok && rank_w == 0) wcnt[w][dg] = cnt_w;
    __syncthreads();
    if (ok) {
      int before = 0;
      for (int ww = 0; ww < w; ww++) before += wcnt[ww][dg];
      const int64_t dst = run[dg] + before + rank_w;
      kout[dst] = key;
      vout[dst] = val;
    }
    __syncthreads();
    // advance the running offsets
    {
      const int d = threadIdx.x;
      int s = 0;
      for (int ww = 0; ww < RS_T / 64; ww++) s += wcnt[ww][d];
      run[d] += s;
    }
    __syncthreads();
  }
}

int psod_radix_sort_u64(psod_ctx *c, int64_t n, uint64_t *keys,
                        uint32_t *vals, int key_bits, const int32_t *d_n) {
  if (n <= 1) return PSOD_OK;
  const int64_t ntile = (n + RS_TILE - 1) / RS_TILE;
  const int64_t hn = ntile * RS_R;
  uint64_t *k2 = (uint64_t *)psod_scratch(c, S_KEYS2, (size_t)n * 8);
  uint32_t *v2 = (uint32_t *)psod_scratch(c, S_VALS2, (size_t)n * 4);
  int32_t *hist = (int32_t *)psod_scratch(c, S_HIST, (size_t)hn * 8 + 64);
  if (!k2 || !v2 || !hist) return PSOD_E_NOMEM;
  int32_t *offs = hist + hn;
  uint64_t *ka = keys, *kb = k2;
  uint32_t *va = vals, *vb = v2;
  int passes = (key_bits + 7) / 8;
  for (int p = 0; p < passes; p++) {
    const int shift = 8 * p;
    k_rs_hist<<<(unsigned)ntile, RS_T, 0, c->stream>>>(ka, n, d_n, shift,
                                                       hist, ntile);
    PSOD_LAUNCH_CHECK();
    PSOD_TRY(psod_scan_i32(c, hist, offs, hn, nullptr));
    k_rs_scatter<<<(unsigned)ntile, RS_T, 0, c->stream>>>(
        ka, va, n, d_n, shift, offs, ntile, kb, vb);
    PSOD_LAUNCH_CHECK();
    uint64_t *tk = ka; ka = kb; kb = tk;
    uint32_t *tv = va; va = vb; vb = tv;
  }
  if (ka != keys) {
    PSOD_HIP(hipMemcpyAsync(keys, ka, (size_t)n * 8, hipMemcpyDeviceToDevice,
                            c->stream));
    PSOD_HIP(hipMemcpyAsync(vals, va, (size_t)n * 4, hipMemcpyDeviceToDevice,
                            c->stream));
  }
  return PSOD_OK;
}
