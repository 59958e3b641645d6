/*
 * gp_hip.hip -- libgp_hip: the C ABI of include/gp_hip.h over hand-written CDNA4 (gfx950) kernels.
 *
 * Device layout (one handle):
 *   every field component is a plane of rows_x = mesh_x+6 rows; a row holds the halo-inclusive y range
 *   0..mesh_y+5 starting at element kYOff, so the first interior cell y = 3 sits on a 64-element boundary;
 *   the row pitch ps is a multiple of 64 elements. phi (N planes) and mu (NC planes) are ping-ponged,
 *   dphi (N planes) holds delta-phi when it is asked for.
 * A time step (default, gp_cells.h) classifies cells by the phase codes of the current state: k_cbulk updates the
 * bulk cells (streaming mu; phi stays e_b) and lists the interface cells and the delta-phi candidates (its blocks
 * past the tiles add the candidates of the boundary ring); k_dlist evaluates delta-phi of the listed candidates and k_ilist the
 * general update of the listed interface cells, both with full wavefronts; k_stats folds the statistics when they
 * are asked for. Every kernel writes the halo images of the boundary cells it updates.
 * GP_PATH=march selects the tiled march (gp_march.h: k_dphi, then k_march near an interface and k_bulk on deep
 * tiles); GP_PATH=twopass the two-pass form (k_phi: delta-phi on x,y in [2, mesh+3]; k_mu: currents and update
 * on the interior), which also runs the smoothing steps and recomputes delta-phi for gp_download.
 * This file holds the C ABI and the host-side handling; the kernels and their launch sequences are templates over
 * (phases, mu fields) in gp_step_impl.h, instantiated once per supported pair by gp_ops.hip.
 * Reference routines replaced: see include/gp_hip.h.
 */
#include <mutex>

#include "gp_internal.h"
#include "gp_adjoint.h"
#include "gp_cells.h"

using namespace gpk;

namespace {
char g_create_error[512];

/* Scratch of gp_invariants_noalloc: the per-block sums of k_invariants, in a device array the library holds for the
 * life of the process (one copy per GPU). Host code never passes `scratch` itself to a kernel (in host code the
 * name stands for a host-side shadow, not a device address): it uses the device address that hipGetSymbolAddress
 * returned for the current GPU, looked up once per GPU when the first handle is created there. */
static __device__ __align__(256) double scratch[GP_NOALLOC_SCRATCH_DOUBLES];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
double *g_scratch_ptr[kMaxDevices];
bool g_scratch_done[kMaxDevices];

/* the device address of `scratch` on GPU `dev` (current), looked up on the first call for that GPU */
int scratch_lookup(int dev, double **out, char *err) {
  if (dev < 0 || dev >= kMaxDevices) return snprintf(err, 512, "device index %d out of range", dev), GP_E_ARG;
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_done[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess)
      return snprintf(err, 512, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e)), GP_E_HIP;
    g_scratch_ptr[dev] = (double *)p;
    g_scratch_done[dev] = true;
  }
  *out = g_scratch_ptr[dev];
  return GP_OK;
}
}  // namespace

int gp_fail(gp_handle *h, int code, const char *fmt, ...) {
  char *dst = h ? h->err : g_create_error;
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(dst, 512, fmt, ap);
  va_end(ap);
  return code;
}

/* the kernel set of (N, NC) in the given precision, or NULL when that pair is not instantiated */
static const GpOps *find_ops(int N, int NC, int precision) {
#define GP_FIND_OPS(a, b) \
  if (N == a && NC == b) return gp_ops_##a##_##b(precision);
  GP_COMBOS(GP_FIND_OPS)
#undef GP_FIND_OPS
  return nullptr;
}

/* Tile grid of the single-pass step: 60-cell y-strips by x-segments of 32 rows (16 when that gives fewer than
 * 4096 tiles, to fill the device on small grids); GP_SEG_LEN overrides the segment length. Each segment
 * recomputes two warm-up rows of delta-phi. */
static void march_plan(gp_handle *h) {
  const int64_t mx = h->p.mesh_x, my = h->p.mesh_y;
  const int64_t nstrips = (my + kMarchOut - 1) / kMarchOut;
  int64_t seg = 32;
  if (nstrips * ((mx + seg - 1) / seg) < 4096) seg = 16;
  const char *e = getenv("GP_SEG_LEN");
  if (e && atoi(e) > 0) seg = atoi(e);
  if (seg > mx) seg = mx;
  h->march_seg = (int)seg;
  h->march_nstrips = (int)nstrips;
  h->march_nsegs = (int)((mx + seg - 1) / seg);
  h->march_nwaves = nstrips * h->march_nsegs;
  /* the cell-classified kernels have almost no per-tile warm-up: 16-row segments (more, smaller tiles) even out
   * the interface work between wavefronts; GP_CELL_SEG overrides (at most kCellSegMax) */
  int64_t cseg = 16;
  const char *ce = getenv("GP_CELL_SEG");
  if (ce && atoi(ce) > 0) cseg = atoi(ce);
  if (cseg > kCellSegMax) cseg = kCellSegMax;
  if (cseg > mx) cseg = mx;
  h->cell_seg = (int)cseg;
  h->cell_nsegs = (int)((mx + cseg - 1) / cseg);
  h->cell_nwaves = nstrips * h->cell_nsegs;
}

/* ------------------------------------------------------------------------------------------------ helpers */

static void stats_reset_host(gp_handle *h, double *s) {
  const int N = h->N, NC = h->NC;
  for (int a = 0; a < N; a++) {
    s[a] = 1.0;      /* phi_max */
    s[N + a] = 0.0;  /* phi_min */
    s[2 * N + 2 * NC + a] = 0.0;
  }
  for (int k = 0; k < NC; k++) {
    s[2 * N + k] = 1.0;       /* mu_max */
    s[2 * N + NC + k] = 1.0;  /* mu_min */
    s[3 * N + 2 * NC + k] = 0.0;
  }
  s[3 * N + 3 * NC] = 0.0;
}

/* halo-inclusive host planes (rows_x*rows_y, y fastest) -> device planes, with the BC maps applied; in GP_FP32C
 * each value is encoded into its FP32 part (dst) and its second part (dst_lo: the FP32 low part of phi, field 0,
 * or the 8-bit extension of mu, field 1) */
static int upload_planes(gp_handle *h, void *dst, void *dst_lo, const double *src, int ncomp, int field) {
  const int64_t rx = h->rx, ry = h->ry, n = rx * ry;
  std::vector<double> tmp(src, src + (size_t)ncomp * n);
  const int64_t hx[6] = {0, 1, 2, rx - 3, rx - 2, rx - 1}, hy[6] = {0, 1, 2, ry - 3, ry - 2, ry - 1};
  for (int c = 0; c < ncomp; c++) {
    double *v = tmp.data() + (size_t)c * n;
    for (int s = 0; s < 6; s++)
      for (int64_t y = 0; y < ry; y++) v[hx[s] * ry + y] = v[(int64_t)h->p.halo_src_x[field][s] * ry + y];
    for (int s = 0; s < 6; s++)
      for (int64_t x = 0; x < rx; x++) v[x * ry + hy[s]] = v[x * ry + h->p.halo_src_y[field][s]];
  }
  const size_t es = esize(h);
  std::vector<char> stage((size_t)ncomp * h->fs * es, 0);
  const size_t les = field == 0 ? sizeof(float) : 1;
  std::vector<char> stage_lo(h->precision == GP_FP32C ? (size_t)ncomp * h->fs * les : 0, 0);
  for (int c = 0; c < ncomp; c++)
    for (int64_t x = 0; x < rx; x++)
      for (int64_t y = 0; y < ry; y++) {
        const double v = tmp[(size_t)c * n + x * ry + y];
        const size_t o = ((size_t)c * h->fs + x * h->ps + y + kYOff);
        if (h->precision == GP_FP64) ((double *)stage.data())[o] = v;
        else if (h->precision == GP_FP32C && field == 0) df_encode(v, ((float *)stage.data())[o], ((float *)stage_lo.data())[o]);
        else if (h->precision == GP_FP32C) c32_encode(v, ((float *)stage.data())[o], ((int8_t *)stage_lo.data())[o]);
        else ((float *)stage.data())[o] = (float)v;
      }
  HIPCHK(h, hipMemcpyAsync(dst, stage.data(), stage.size(), hipMemcpyHostToDevice, h->stream));
  if (h->precision == GP_FP32C)
    HIPCHK(h, hipMemcpyAsync(dst_lo, stage_lo.data(), stage_lo.size(), hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return GP_OK;
}

/* device planes -> halo-inclusive host planes. es: element size of the planes (8: FP64, 4: FP32); src_lo: the
 * GP_FP32C second parts of FP32 planes (decoded into the values: FP32 low parts of phi, field 0, or 8-bit
 * extensions of mu, field 1), or NULL */
static int download_planes(gp_handle *h, double *dst, const void *src, const void *src_lo, int field, int ncomp,
                           size_t es) {
  const int64_t rx = h->rx, ry = h->ry, n = rx * ry;
  std::vector<char> stage((size_t)ncomp * h->fs * es);
  const size_t les = field == 0 ? sizeof(float) : 1;
  std::vector<char> stage_lo(src_lo ? (size_t)ncomp * h->fs * les : 0);
  HIPCHK(h, hipMemcpyAsync(stage.data(), src, stage.size(), hipMemcpyDeviceToHost, h->stream));
  if (src_lo) HIPCHK(h, hipMemcpyAsync(stage_lo.data(), src_lo, stage_lo.size(), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  for (int c = 0; c < ncomp; c++)
    for (int64_t x = 0; x < rx; x++)
      for (int64_t y = 0; y < ry; y++) {
        const size_t o = ((size_t)c * h->fs + x * h->ps + y + kYOff);
        double v;
        if (es == sizeof(double)) v = ((double *)stage.data())[o];
        else if (src_lo && field == 0) v = df_decode(((float *)stage.data())[o], ((float *)stage_lo.data())[o]);
        else if (src_lo) v = c32_decode(((float *)stage.data())[o], ((int8_t *)stage_lo.data())[o]);
        else v = (double)((float *)stage.data())[o];
        dst[(size_t)c * n + x * ry + y] = v;
      }
  return GP_OK;
}

static int validate(const gp_params *p, char *err) {
  if (p->nphases < 2 || p->nphases > GP_MAXP) return snprintf(err, 512, "nphases %d outside [2,%d]", p->nphases, GP_MAXP), GP_E_ARG;
  if (p->ncomp < 2 || p->ncomp - 1 > GP_MAXNC) return snprintf(err, 512, "ncomp %d outside [2,%d]", p->ncomp, GP_MAXNC + 1), GP_E_ARG;
  if (p->mesh_x < 4 || p->mesh_y < 4) return snprintf(err, 512, "mesh must be at least 4x4"), GP_E_ARG;
  if (p->mesh_x > (1 << 20) || p->mesh_y > (1 << 20)) return snprintf(err, 512, "mesh too large"), GP_E_ARG;
  if (p->fn_w != 1 && p->fn_w != 2) return snprintf(err, 512, "fn_w must be 1 (obstacle) or 2 (well)"), GP_E_ARG;
  if (p->fn_aniso != 0 && p->fn_aniso != 1) return snprintf(err, 512, "fn_aniso must be 0 or 1"), GP_E_ARG;
  if (p->fn_aniso == 1 && p->fold != 4) return snprintf(err, 512, "anisotropy type %d not supported (4 only)", p->fold), GP_E_ARG;
  if (!(p->dx > 0 && p->dy > 0 && p->dt > 0 && p->epsilon > 0)) return snprintf(err, 512, "dx, dy, dt, epsilon must be > 0"), GP_E_ARG;
  if (p->solve < 0 || p->solve > 3) return snprintf(err, 512, "unknown solve mode %d", p->solve), GP_E_ARG;
  if (p->solve == GP_SOLVE_BINARY && p->ncomp != 2) return snprintf(err, 512, "BINARY solve needs ncomp = 2"), GP_E_ARG;
  if (p->solve == GP_SOLVE_TERNARY && p->ncomp != 3) return snprintf(err, 512, "TERNARY solve needs ncomp = 3"), GP_E_ARG;
  const int64_t ex[2] = {p->mesh_x + 2, p->mesh_y + 2};
  for (int f = 0; f < 2; f++)
    for (int k = 0; k < 6; k++) {
      if (p->halo_src_x[f][k] < 3 || p->halo_src_x[f][k] > ex[0] || p->halo_src_y[f][k] < 3 || p->halo_src_y[f][k] > ex[1])
        return snprintf(err, 512, "halo map of field %d entry %d points outside the interior", f, k), GP_E_ARG;
    }
  return GP_OK;
}

/* ------------------------------------------------------------------------------------------------ adjoint */

/* the current state into / from a state buffer laid out phi planes then mu planes */
/* (a GP_FP32C state buffer holds the extensions after the planes) */
static int state_save(gp_handle *h, double *dst) {
  const size_t es = esize(h), pb = (size_t)h->N * h->fs * es, mb = (size_t)h->NC * h->fs * es;
  HIPCHK(h, hipMemcpyAsync(dst, h->phi[h->cur], pb, hipMemcpyDeviceToDevice, h->stream));
  HIPCHK(h, hipMemcpyAsync((char *)dst + pb, h->mu[h->cur], mb, hipMemcpyDeviceToDevice, h->stream));
  if (h->precision == GP_FP32C)
    HIPCHK(h, hipMemcpyAsync((char *)dst + pb + mb, h->lo[h->cur], c32_lo_bytes(h), hipMemcpyDeviceToDevice,
                             h->stream));
  return GP_OK;
}
static int state_load(gp_handle *h, const double *src) {
  const size_t es = esize(h), pb = (size_t)h->N * h->fs * es, mb = (size_t)h->NC * h->fs * es;
  HIPCHK(h, hipMemcpyAsync(h->phi[h->cur], src, pb, hipMemcpyDeviceToDevice, h->stream));
  HIPCHK(h, hipMemcpyAsync(h->mu[h->cur], (const char *)src + pb, mb, hipMemcpyDeviceToDevice, h->stream));
  if (h->precision == GP_FP32C)
    HIPCHK(h, hipMemcpyAsync(h->lo[h->cur], (const char *)src + pb + mb, c32_lo_bytes(h), hipMemcpyDeviceToDevice,
                             h->stream));
  h->flags_valid = false;
  h->codes_valid = false;
  h->prev_valid = false;
  return GP_OK;
}

/* halo-inclusive host planes (rows_x*rows_y, y fastest) <-> device planes, without boundary copies: one strided
 * copy (the planes are rows_x rows apart on both sides) */
static int planes_to_device(gp_handle *h, double *dst, const double *src, int ncomp) {
  HIPCHK(h, hipMemcpy2DAsync(dst + kYOff, h->ps * sizeof(double), src, h->ry * sizeof(double), h->ry * sizeof(double),
                             (size_t)ncomp * h->rx, hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return GP_OK;
}
static int planes_to_host(gp_handle *h, double *dst, const double *src, int ncomp) {
  HIPCHK(h, hipMemcpy2DAsync(dst, h->ry * sizeof(double), src + kYOff, h->ps * sizeof(double), h->ry * sizeof(double),
                             (size_t)ncomp * h->rx, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return GP_OK;
}

/* GP_ADJOINT_CHECK: the smallest index of a non-finite value of g[0, n) into out[0] (atomicMin; stays at ~0 when
 * there is none) and the largest |g| among the finite ones, as the bits of a non-negative double, into out[1] */
static __global__ __launch_bounds__(256) void k_adj_scan(const double *__restrict__ g, int64_t n,
                                                         unsigned long long *out) {
  unsigned long long bad = ~0ull, mx = 0ull;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
    const double v = g[i];
    if (!isfinite(v)) bad = (unsigned long long)i < bad ? (unsigned long long)i : bad;
    else {
      const unsigned long long b = (unsigned long long)__double_as_longlong(fabs(v));
      mx = b > mx ? b : mx;
    }
  }
  if (bad != ~0ull) atomicMin(&out[0], bad);
  if (mx) atomicMax(&out[1], mx);
}

/* after transposing forward step t (state t -> t + 1): stop with GP_E_NAN at the first non-finite adjoint or theta
 * gradient, naming the cell and the forward state there; every `every` steps print the largest |adjoint| */
static int adj_check(gp_handle *h, int64_t t, int64_t every, const double *phi64, const double *mu64) {
  Adjoint *A = h->adj;
  const int nth = theta_count(h->N, h->NC);
  if (!A->scan) HIPCHK(h, hipMalloc((void **)&A->scan, 4 * sizeof(unsigned long long)));
  const unsigned long long init[4] = {~0ull, 0ull, ~0ull, 0ull};
  HIPCHK(h, hipMemcpyAsync(A->scan, init, sizeof init, hipMemcpyHostToDevice, h->stream));
  const int64_t se = (int64_t)state_elems(h);
  hipLaunchKernelGGL(k_adj_scan, dim3(1024), dim3(256), 0, h->stream, (const double *)A->g[1], se, A->scan);
  hipLaunchKernelGGL(k_adj_scan, dim3(1), dim3(256), 0, h->stream, (const double *)A->gth, (int64_t)nth, A->scan + 2);
  HIPCHK(h, hipGetLastError());
  unsigned long long r[4];
  HIPCHK(h, hipMemcpyAsync(r, A->scan, sizeof r, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  double gmax;
  memcpy(&gmax, &r[1], sizeof gmax);
  if (every > 0 && t % every == 0) fprintf(stderr, "gp adjoint check: step %lld max |adjoint| %.6e\n", (long long)t, gmax);
  if (r[0] == ~0ull && r[2] == ~0ull) return GP_OK;
  if (r[0] == ~0ull)
    return gp_fail(h, GP_E_NAN, "adjoint check: theta gradient %llu non-finite after transposing step %lld",
                   (unsigned long long)r[2], (long long)t);
  const int64_t comp = (int64_t)(r[0] / (unsigned long long)h->fs), c = (int64_t)(r[0] % (unsigned long long)h->fs);
  const int64_t x = (c - kYOff) / h->ps, y = (c - kYOff) % h->ps;
  double ph[P] = {0}, mv[NCM] = {0};
  for (int a = 0; a < h->N; a++)
    HIPCHK(h, hipMemcpy(&ph[a], phi64 + (size_t)a * h->fs + c, sizeof(double), hipMemcpyDeviceToHost));
  for (int k = 0; k < h->NC; k++)
    HIPCHK(h, hipMemcpy(&mv[k], mu64 + (size_t)k * h->fs + c, sizeof(double), hipMemcpyDeviceToHost));
  return gp_fail(h, GP_E_NAN,
                 "adjoint check: non-finite adjoint of %s %lld at cell (x %lld, y %lld) after transposing step %lld "
                 "(largest finite |adjoint| %.3e); state there: phi %.17g %.17g %.17g %.17g, mu %.17g %.17g %.17g",
                 comp < h->N ? "phi" : "mu", (long long)(comp < h->N ? comp : comp - h->N), (long long)x,
                 (long long)y, (long long)t, gmax, ph[0], ph[1], ph[2], ph[3], mv[0], mv[1], mv[2]);
}

/* fold the halo adjoints of A->g[1] into their source cells */
static int adj_fold(gp_handle *h) {
  Adjoint *A = h->adj;
  const dim3 gall((unsigned)((h->ry + 63) / 64), (unsigned)((h->rx + 3) / 4));
  hipLaunchKernelGGL(k_adj_bc, dim3(gall.x, gall.y, 2), dim3(256), 0, h->stream, A->g[1], h->N, h->NC, h->maps, h->fs,
                     h->ps, h->rx, h->ry);
  HIPCHK(h, hipGetLastError());
  return GP_OK;
}

/* ------------------------------------------------------------------------------------------------ frame */

/* temperature of allocated row 0 and the gradient per row, as the serial driver forms them before the loop (t = 0,
 * GPs/microsim_gp.c:77-80) and after step t (:108-111) */
static double frame_bottom(const gp_handle *h, int64_t t, double *grad) {
  const gp_params *p = &h->p;
  const double G = (p->tg_deltaT) * p->dy / (p->tg_distance);
  double base;
  if (t == 0) base = (p->tg_offset / p->dy) - (double)h->shift_off;
  else base = (p->tg_offset / p->dy) - (double)h->shift_off + ((p->tg_velocity / p->dy) * ((double)t * p->dt));
  *grad = G;
  return p->tg_base_temp - base * G;
}

/* T of every allocated row before the first step: the t = 0 gradient, then the boundary copies of T that run
 * before the loop (GPs/microsim_gp.c:81-84) */
static int frame_reset(gp_handle *h) {
  h->shift_off = 0;
  if (!h->p.tempgrady) return GP_OK;
  double G;
  const double tb = frame_bottom(h, 0, &G);
  std::vector<double> T((size_t)h->ry);
  for (int64_t y = 0; y < h->ry; y++) T[y] = tb + (double)y * G;
  for (int k = 0; k < 6; k++) T[h->tp.hy[k]] = T[h->tp.srcT[k]];
  HIPCHK(h, hipMemcpyAsync(h->trow, T.data(), T.size() * sizeof(double), hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return GP_OK;
}

/* after step t: the new temperature gradient, and every 100 steps the moving frame */
static int frame_after_step(gp_handle *h, int64_t t) {
  if (h->p.tempgrady) {
    double G;
    const double tb = frame_bottom(h, t, &G);
    hipLaunchKernelGGL(k_trow_set, dim3((unsigned)((h->ry + 255) / 256)), dim3(256), 0, h->stream, h->trow, h->ry, tb,
                       G);
    HIPCHK(h, hipGetLastError());
  }
  if (!h->p.shift || t % 100 != 0) return GP_OK;
  const int N = h->N, NC = h->NC, L = N - 1;
  const size_t es = esize(h);
  HIPCHK(h, hipMemsetAsync(h->front, 0, sizeof(int), h->stream));
  const dim3 gf((unsigned)((h->ry - 1 + 63) / 64), (unsigned)((h->p.mesh_x + 3) / 4));
  if (h->precision == GP_FP64)
    hipLaunchKernelGGL((k_front<double>), gf, dim3(256), 0, h->stream,
                       (const double *)h->phi[h->cur] + (size_t)L * h->fs, h->front, h->ps, h->p.mesh_x, h->ry);
  else
    hipLaunchKernelGGL((k_front<float>), gf, dim3(256), 0, h->stream,
                       (const float *)h->phi[h->cur] + (size_t)L * h->fs, h->front, h->ps, h->p.mesh_x, h->ry);
  HIPCHK(h, hipGetLastError());
  int pos = 0;
  HIPCHK(h, hipMemcpyAsync(&pos, h->front, sizeof(int), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  if (pos <= h->p.shiftj) return GP_OK;
  const int64_t sh = pos - h->p.shiftj;
  FillValues fp = {}, fm = {};
  fp.v[L] = 1.0;
  for (int k = 0; k < NC; k++) fm.v[k] = h->p.mu_fill[k];
  const dim3 ga((unsigned)((h->ry + 63) / 64), (unsigned)((h->rx + 3) / 4));
  if (es == sizeof(double)) {
    hipLaunchKernelGGL((k_shift<double>), ga, dim3(256), 0, h->stream, (const double *)h->phi[h->cur],
                       (double *)h->phi[1 - h->cur], N, fp, sh, h->fs, h->ps, h->rx, h->ry);
    hipLaunchKernelGGL((k_shift<double>), ga, dim3(256), 0, h->stream, (const double *)h->mu[h->cur],
                       (double *)h->mu[1 - h->cur], NC, fm, sh, h->fs, h->ps, h->rx, h->ry);
  } else {
    hipLaunchKernelGGL((k_shift<float>), ga, dim3(256), 0, h->stream, (const float *)h->phi[h->cur],
                       (float *)h->phi[1 - h->cur], N, fp, sh, h->fs, h->ps, h->rx, h->ry);
    hipLaunchKernelGGL((k_shift<float>), ga, dim3(256), 0, h->stream, (const float *)h->mu[h->cur],
                       (float *)h->mu[1 - h->cur], NC, fm, sh, h->fs, h->ps, h->rx, h->ry);
  }
  HIPCHK(h, hipGetLastError());
  h->cur = 1 - h->cur;
  if (h->p.tempgrady) {
    /* T moves with the rows; the top rows keep their own */
    std::vector<double> T((size_t)h->ry);
    HIPCHK(h, hipMemcpyAsync(T.data(), h->trow, T.size() * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    for (int64_t y = 0; y + sh <= h->ry - 1; y++) T[y] = T[y + sh];
    HIPCHK(h, hipMemcpyAsync(h->trow, T.data(), T.size() * sizeof(double), hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
  }
  h->shift_off += sh;
  h->codes_valid = false;
  h->flags_valid = false;
  h->prev_valid = false;
  return GP_OK;
}

/* ------------------------------------------------------------------------------------------------ C ABI */

extern "C" {

const char *gp_version(void) { return "gp_hip 0.1 gfx950"; }

const char *gp_last_error(const gp_handle *h) { return h ? h->err : g_create_error; }

int gp_create(const gp_params *params, int precision, gp_handle **out) {
  if (!params || !out) return gp_fail(nullptr, GP_E_ARG, "null argument");
  *out = nullptr;
  int rc = validate(params, g_create_error);
  if (rc) return rc;
  if (precision != GP_FP64 && precision != GP_FP32 && precision != GP_FP32C)
    return gp_fail(nullptr, GP_E_ARG, "unknown precision %d", precision);
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return gp_fail(nullptr, GP_E_NODEV, "no HIP device");
  gp_handle *h = new gp_handle();
  memset(h->err, 0, sizeof h->err);
  h->p = *params;
  h->precision = precision;
  h->N = params->nphases;
  h->NC = params->ncomp - 1;
  h->rx = params->mesh_x + 6;
  h->ry = params->mesh_y + 6;
  h->ps = ((h->ry + kYOff + 63) / 64) * 64;
  h->fs = h->rx * h->ps;
  h->cur = 0;
  h->steps = 0;
  h->ops = find_ops(h->N, h->NC, precision);
  if (!h->ops) {
    delete h;
    return gp_fail(nullptr, GP_E_ARG, "no kernel instantiated for %d phases / %d components", params->nphases, params->ncomp);
  }
  const int64_t hx[6] = {0, 1, 2, h->rx - 3, h->rx - 2, h->rx - 1}, hy[6] = {0, 1, 2, h->ry - 3, h->ry - 2, h->ry - 1};
  for (int f = 0; f < 2; f++)
    for (int k = 0; k < 6; k++) {
      h->maps.sx[f][k] = params->halo_src_x[f][k];
      h->maps.sy[f][k] = params->halo_src_y[f][k];
    }
  for (int k = 0; k < 6; k++) {
    h->maps.hx[k] = (int32_t)hx[k];
    h->maps.hy[k] = (int32_t)hy[k];
  }
  if (hipGetDevice(&h->device) != hipSuccess) {
    delete h;
    return gp_fail(nullptr, GP_E_NODEV, "hipGetDevice failed");
  }
  {
    /* device address of gp_invariants_noalloc's scratch on this GPU (looked up once per GPU) */
    double *sp = nullptr;
    rc = scratch_lookup(h->device, &sp, g_create_error);
    if (rc) {
      delete h;
      return rc;
    }
  }
  h->nblocks_mu = ((params->mesh_y + kBX - 1) / kBX) * ((params->mesh_x + kBY - 1) / kBY);
  march_plan(h);
  h->nparts_max = std::max(h->nblocks_mu, std::max(2 * h->march_nwaves, h->cell_nwaves + 4 * (int64_t)kListBlocks));
  const char *path = getenv("GP_PATH");
  h->use_march = !(path && strcmp(path, "twopass") == 0);
  /* the cell lists and k_cbulk's row offsets are 32-bit: grids beyond that bound (cells_fit, about 46000^2 cells)
   * take the tiled march, which gives the same results */
  h->use_cells = h->use_march && !(path && strcmp(path, "march") == 0) && cells_fit(h->rx, h->ps);
  if ((params->tempgrady || params->shift) && precision == GP_FP32C) {
    delete h;
    return gp_fail(nullptr, GP_E_ARG, "non-isothermal runs and the moving frame run in GP_FP64 or GP_FP32");
  }
  if (params->tempgrady) h->use_march = h->use_cells = false; /* the two-pass kernels carry the row temperatures */
  if (precision == GP_FP32C && !h->use_cells) {
    delete h;
    return gp_fail(nullptr, GP_E_ARG,
                   "the compensated FP32 mode runs the cell-classified step only (GP_PATH selects another path, or the "
                   "grid has 2^31 cells or more)");
  }
  h->prev_valid = false;
  h->fcur = 0;
  h->flags_valid = false;
  h->codes_valid = false;
  const size_t es = precision == GP_FP64 ? 8 : 4;
  const size_t pn = (size_t)h->N * h->fs * es, mn = (size_t)h->NC * h->fs * es;
  hipError_t e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
  bool ok = e == hipSuccess;
  ok = ok && hipMalloc(&h->phi[0], pn) == hipSuccess && hipMalloc(&h->phi[1], pn) == hipSuccess;
  ok = ok && hipMalloc(&h->mu[0], mn) == hipSuccess && hipMalloc(&h->mu[1], mn) == hipSuccess;
  ok = ok && hipMalloc(&h->dphi, (size_t)h->N * h->fs * dphi_esize(h)) == hipSuccess;
  if (params->tempgrady) {
    const size_t rt = precision == GP_FP64 ? sizeof(RowThermo<double>) : sizeof(RowThermo<float>);
    ok = ok && hipMalloc((void **)&h->trow, (size_t)h->ry * sizeof(double)) == hipSuccess &&
         hipMalloc(&h->thp, (size_t)h->ry * rt) == hipSuccess && hipMalloc(&h->thc, (size_t)h->ry * rt) == hipSuccess;
  }
  if (params->shift) ok = ok && hipMalloc((void **)&h->front, sizeof(int)) == hipSuccess;
  if (precision == GP_FP32C)
    ok = ok && hipMalloc((void **)&h->lo[0], c32_lo_bytes(h)) == hipSuccess &&
         hipMalloc((void **)&h->lo[1], c32_lo_bytes(h)) == hipSuccess &&
         hipMalloc(&h->model64, sizeof(Model<double>)) == hipSuccess;
  ok = ok && hipMalloc((void **)&h->partials, sizeof(double) * kNStatMax * h->nparts_max) == hipSuccess;
  ok = ok && hipMalloc((void **)&h->stats_dev, sizeof(double) * kNStatMax) == hipSuccess;
  ok = ok && hipMalloc((void **)&h->inv_dev, sizeof(double) * (GP_MAXNC + 1)) == hipSuccess;
  ok = ok && hipMalloc((void **)&h->flags[0], sizeof(int) * h->march_nwaves) == hipSuccess;
  ok = ok && hipMalloc((void **)&h->flags[1], sizeof(int) * h->march_nwaves) == hipSuccess;
  ok = ok && hipMalloc((void **)&h->clist, (size_t)h->fs * sizeof(int32_t)) == hipSuccess &&
       hipMalloc((void **)&h->ilist, (size_t)h->fs * sizeof(int32_t)) == hipSuccess &&
       hipMalloc((void **)&h->lcount, 4 * sizeof(int)) == hipSuccess;
  ok = ok && hipMalloc((void **)&h->codes[0], (size_t)h->rx * h->ps) == hipSuccess &&
       hipMalloc((void **)&h->codes[1], (size_t)h->rx * h->ps) == hipSuccess;
  h->ccur = 0;
  const size_t msize = precision == GP_FP64 ? sizeof(Model<double>) : sizeof(Model<float>);
  ok = ok && hipMalloc(&h->model, msize) == hipSuccess;
  if (!ok) {
    gp_destroy(h);
    return gp_fail(nullptr, GP_E_OOM, "device allocation failed (%zu bytes per phi copy)", pn);
  }
  ok = hipMemsetAsync(h->phi[0], 0, pn, h->stream) == hipSuccess && hipMemsetAsync(h->phi[1], 0, pn, h->stream) == hipSuccess &&
       hipMemsetAsync(h->mu[0], 0, mn, h->stream) == hipSuccess && hipMemsetAsync(h->mu[1], 0, mn, h->stream) == hipSuccess &&
       hipMemsetAsync(h->dphi, 0, (size_t)h->N * h->fs * dphi_esize(h), h->stream) == hipSuccess &&
       hipMemsetAsync(h->lcount, 0, 4 * sizeof(int), h->stream) == hipSuccess;
  h->lsel = 0;
  if (precision == GP_FP64) {
    Model<double> m;
    make_model(params, &m);
    h->pow2 = m.pow2 != 0;
    ok = ok && hipMemcpyAsync(h->model, &m, sizeof m, hipMemcpyHostToDevice, h->stream) == hipSuccess;
    ok = ok && hipStreamSynchronize(h->stream) == hipSuccess;
  } else {
    Model<float> m;
    make_model(params, &m);
    h->pow2 = m.pow2 != 0;
    ok = ok && hipMemcpyAsync(h->model, &m, sizeof m, hipMemcpyHostToDevice, h->stream) == hipSuccess;
    if (precision == GP_FP32C) {
      Model<double> m64;
      make_model(params, &m64);
      ok = ok && hipMemsetAsync(h->lo[0], 0, c32_lo_bytes(h), h->stream) == hipSuccess &&
           hipMemsetAsync(h->lo[1], 0, c32_lo_bytes(h), h->stream) == hipSuccess &&
           hipMemcpyAsync(h->model64, &m64, sizeof m64, hipMemcpyHostToDevice, h->stream) == hipSuccess;
      ok = ok && hipStreamSynchronize(h->stream) == hipSuccess;
    }
    ok = ok && hipStreamSynchronize(h->stream) == hipSuccess;
  }
  if (!ok) {
    gp_destroy(h);
    return gp_fail(nullptr, GP_E_HIP, "device initialisation failed");
  }
  if (params->tempgrady) {
    /* the row coefficients' parameters (gp_frame.h k_thermo) */
    ThermoParams &tp = h->tp;
    tp.N = h->N;
    tp.NC = h->NC;
    tp.Teq = params->Teq;
    for (int k = 0; k < 6; k++) {
      tp.hy[k] = h->maps.hy[k];
      tp.srcT[k] = params->halo_src_T_y[k];
    }
    for (int a = 0; a < GP_MAXP; a++) {
      for (int i = 0; i < GP_MAXNC; i++) {
        tp.Beq[a][i] = params->Beq[a * GP_MAXNC + i];
        tp.dBbdT[a][i] = params->dBbdT[a * GP_MAXNC + i];
        tp.ceq_sol[a][i] = params->ceq_sol[a * GP_MAXNC + i];
        tp.ceq_liq[a][i] = params->ceq_liq[a * GP_MAXNC + i];
        tp.delta_c[a][i] = params->delta_c[a * GP_MAXNC + i];
        for (int j = 0; j < GP_MAXNC; j++) tp.A[a][i][j] = params->A[(a * GP_MAXNC + i) * GP_MAXNC + j];
      }
      tp.delta_T_sol[a] = params->delta_T_sol[a];
      tp.delta_T_liq[a] = params->delta_T_liq[a];
    }
  }
  double s[kNStatMax] = {0};
  stats_reset_host(h, s);
  if (hipMemcpy(h->stats_dev, s, sizeof s, hipMemcpyHostToDevice) != hipSuccess) {
    gp_destroy(h);
    return gp_fail(nullptr, GP_E_HIP, "device initialisation failed");
  }
  *out = h;
  return GP_OK;
}

void gp_destroy(gp_handle *h) {
  if (!h) return;
  /* teardown: errors here have no caller to report to, so they are ignored explicitly */
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  for (auto &r : h->prof) {
    (void)hipEventDestroy(r.second.first);
    (void)hipEventDestroy(r.second.second);
  }
  h->prof.clear();
  delete h->adj;
  void *bufs[] = {h->phi[0], h->phi[1], h->mu[0],    h->mu[1],    h->dphi,   h->partials, h->stats_dev,
                  h->model,  h->flags[0], h->flags[1], h->inv_dev, h->codes[0], h->codes[1], h->clist,
                  h->ilist,  h->lcount, h->lo[0],    h->lo[1],    h->model64, h->wide[0],  h->wide[1],
                  h->trow,   h->thp,    h->thc,      h->front};
  for (void *b : bufs)
    if (b) (void)hipFree(b);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
}

int gp_upload(gp_handle *h, const double *phi, const double *mu) {
  if (!h || !phi || !mu) return h ? gp_fail(h, GP_E_ARG, "null argument") : GP_E_ARG;
  const bool c = h->precision == GP_FP32C;
  int rc = upload_planes(h, h->phi[h->cur], c ? (void *)c32_plo(h, h->cur) : nullptr, phi, h->N, 0);
  if (rc) return rc;
  rc = upload_planes(h, h->mu[h->cur], c ? (void *)c32_mlo(h, h->cur) : nullptr, mu, h->NC, 1);
  if (rc) return rc;
  /* the other copy starts identical so its halo rows are defined */
  HIPCHK(h, hipMemcpyAsync(h->phi[1 - h->cur], h->phi[h->cur], (size_t)h->N * h->fs * esize(h), hipMemcpyDeviceToDevice, h->stream));
  HIPCHK(h, hipMemcpyAsync(h->mu[1 - h->cur], h->mu[h->cur], (size_t)h->NC * h->fs * esize(h), hipMemcpyDeviceToDevice, h->stream));
  if (c) HIPCHK(h, hipMemcpyAsync(h->lo[1 - h->cur], h->lo[h->cur], c32_lo_bytes(h), hipMemcpyDeviceToDevice, h->stream));
  HIPCHK(h, hipMemsetAsync(h->dphi, 0, (size_t)h->N * h->fs * dphi_esize(h), h->stream));
  double s[kNStatMax] = {0};
  stats_reset_host(h, s);
  HIPCHK(h, hipMemcpyAsync(h->stats_dev, s, sizeof s, hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  h->steps = 0;
  h->prev_valid = false;
  h->flags_valid = false;
  h->codes_valid = false;
  return frame_reset(h);
}

int gp_download(gp_handle *h, double *phi, double *mu, double *dphi) {
  if (!h) return GP_E_ARG;
  int rc;
  const bool c = h->precision == GP_FP32C;
  const void *plo = c ? (const void *)c32_plo(h, h->cur) : nullptr, *mlo = c ? (const void *)c32_mlo(h, h->cur) : nullptr;
  if (phi && (rc = download_planes(h, phi, h->phi[h->cur], plo, 0, h->N, esize(h)))) return rc;
  if (mu && (rc = download_planes(h, mu, h->mu[h->cur], mlo, 1, h->NC, esize(h)))) return rc;
  if (dphi) {
    if (h->prev_valid) {
      /* single-pass steps do not store delta-phi: recompute the last step's from the previous state */
      if ((rc = h->ops->step(h, MODE_DPHI_PREV))) return rc;
    }
    if ((rc = download_planes(h, dphi, h->dphi, nullptr, 0, h->N, dphi_esize(h)))) return rc;
  }
  return GP_OK;
}

int gp_upload_fields(gp_handle *h, const gp_cell *cells, int64_t ncells) {
  if (!h || !cells) return h ? gp_fail(h, GP_E_ARG, "null argument") : GP_E_ARG;
  const int64_t n = h->rx * h->ry;
  if (ncells != n) return gp_fail(h, GP_E_ARG, "ncells %lld != rows_x*rows_y %lld", (long long)ncells, (long long)n);
  std::vector<double> phi((size_t)h->N * n), mu((size_t)h->NC * n);
  for (int64_t i = 0; i < n; i++) {
    for (int a = 0; a < h->N; a++) phi[(size_t)a * n + i] = cells[i].phia[a];
    for (int k = 0; k < h->NC; k++) mu[(size_t)k * n + i] = cells[i].compi[k];
  }
  return gp_upload(h, phi.data(), mu.data());
}

int gp_download_fields(gp_handle *h, gp_cell *cells, int64_t ncells) {
  if (!h || !cells) return h ? gp_fail(h, GP_E_ARG, "null argument") : GP_E_ARG;
  const int64_t n = h->rx * h->ry;
  if (ncells != n) return gp_fail(h, GP_E_ARG, "ncells %lld != rows_x*rows_y %lld", (long long)ncells, (long long)n);
  std::vector<double> phi((size_t)h->N * n), mu((size_t)h->NC * n), dphi((size_t)h->N * n), T((size_t)h->ry);
  int rc = gp_download(h, phi.data(), mu.data(), dphi.data());
  if (rc || (rc = gp_get_frame(h, nullptr, T.data()))) return rc;
  for (int64_t i = 0; i < n; i++) {
    cells[i].temperature = T[(size_t)(i % h->ry)];
    for (int a = 0; a < h->N; a++) {
      cells[i].phia[a] = phi[(size_t)a * n + i];
      if (cells[i].deltaphi) cells[i].deltaphi[a] = dphi[(size_t)a * n + i];
    }
    for (int k = 0; k < h->NC; k++) cells[i].compi[k] = mu[(size_t)k * n + i];
  }
  return GP_OK;
}

int gp_smooth(gp_handle *h, int64_t nsteps) {
  if (!h || nsteps < 0) return h ? gp_fail(h, GP_E_ARG, "bad nsteps") : GP_E_ARG;
  step_fn f = h->ops->step;
  {
    ProfScope ps(h, GP_KERNEL_CALL);
    for (int64_t t = 0; t < nsteps; t++) {
      int rc = f(h, MODE_SMOOTH);
      if (rc) return rc;
    }
  }
  if (nsteps > 0) h->prev_valid = false; /* dphi now holds the smoothing increment */
  return GP_OK;
}

int gp_get_stats(gp_handle *h, gp_stats *st) {
  if (!h || !st) return h ? gp_fail(h, GP_E_ARG, "null argument") : GP_E_ARG;
  double s[kNStatMax], inv[GP_MAXNC + 1];
  inv_fn fi = h->ops->invariants;
  int rc = fi(h);
  if (rc) return rc;
  HIPCHK(h, hipMemcpyAsync(s, h->stats_dev, sizeof s, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipMemcpyAsync(inv, h->inv_dev, sizeof(double) * (h->NC + 1), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  memset(st, 0, sizeof *st);
  const int N = h->N, NC = h->NC;
  for (int a = 0; a < N; a++) {
    st->phi_max[a] = s[a];
    st->phi_min[a] = s[N + a];
    st->rel_change_phi[a] = s[2 * N + 2 * NC + a];
  }
  for (int k = 0; k < NC; k++) {
    st->mu_max[k] = s[2 * N + k];
    st->mu_min[k] = s[2 * N + NC + k];
    st->rel_change_mu[k] = s[3 * N + 2 * NC + k];
  }
  st->nonfinite = (int64_t)s[3 * N + 3 * NC];
  for (int k = 0; k < NC; k++) st->solute[k] = inv[k];
  st->energy = inv[NC];
  st->step = h->steps;
  return GP_OK;
}

int gp_get_frame(gp_handle *h, int64_t *shift_offset, double *T_rows) {
  if (!h) return GP_E_ARG;
  if (shift_offset) *shift_offset = h->shift_off;
  if (T_rows) {
    if (h->p.tempgrady) {
      HIPCHK(h, hipMemcpyAsync(T_rows, h->trow, (size_t)h->ry * sizeof(double), hipMemcpyDeviceToHost, h->stream));
      HIPCHK(h, hipStreamSynchronize(h->stream));
    } else {
      for (int64_t y = 0; y < h->ry; y++) T_rows[y] = h->p.T;
    }
  }
  return GP_OK;
}

int gp_invariants_noalloc(gp_handle *h, const void *phi, const void *mu, double *out, void *stream) {
  if (!h || !phi || !mu || !out) return h ? gp_fail(h, GP_E_ARG, "null argument") : GP_E_ARG;
  const int64_t mx = h->p.mesh_x, my = h->p.mesh_y;
  /* the block grid of k_invariants (one block per 64 y-cells by 4 x-rows of the interior) and its block sums */
  const int64_t nblocks = ((my + 63) / 64) * ((mx + 3) / 4);
  const int64_t need = (int64_t)(h->NC + 1) * nblocks;
  if (need > (int64_t)GP_NOALLOC_SCRATCH_DOUBLES)
    return gp_fail(h, GP_E_ARG,
                   "gp_invariants_noalloc: a %lldx%lld grid needs %lld scratch doubles, more than the %lld the library "
                   "holds (GP_NOALLOC_SCRATCH_DOUBLES)",
                   (long long)mx, (long long)my, (long long)need, (long long)GP_NOALLOC_SCRATCH_DOUBLES);
  int dev = -1;
  HIPCHK(h, hipGetDevice(&dev));
  if (dev != h->device)
    return gp_fail(h, GP_E_ARG, "gp_invariants_noalloc: the handle lives on device %d, device %d is current",
                   h->device, dev);
  double *sp = nullptr;
  int rc = scratch_lookup(dev, &sp, h->err);
  if (rc) return rc;
  /* the caller's state is in the reference's plain layout: plane = rows_x * rows_y, row pitch rows_y, no offset */
  return h->ops->invariants_on(h, phi, mu, h->rx * h->ry, h->ry, 0, sp, out,
                               stream ? (hipStream_t)stream : h->stream);
}

int gp_step(gp_handle *h, int64_t nsteps, gp_stats *stats) {
  if (!h || nsteps < 0) return h ? gp_fail(h, GP_E_ARG, "bad nsteps") : GP_E_ARG;
  step_fn f = h->ops->step;
  {
    ProfScope ps(h, GP_KERNEL_CALL);
    const bool framed = h->p.tempgrady || h->p.shift;
    for (int64_t t = 0; t < nsteps; t++) {
      int rc = f(h, h->use_march ? (stats ? MODE_MARCH_STATS : MODE_MARCH) : MODE_TWOPASS);
      if (rc) return rc;
      h->steps++;
      h->prev_valid = h->use_march;
      if (framed && (rc = frame_after_step(h, h->steps))) return rc;
    }
  }
  if (stats) {
    int rc = gp_get_stats(h, stats);
    if (rc) return rc;
    if (stats->nonfinite) return gp_fail(h, GP_E_NAN, "%lld non-finite values after step %lld", (long long)stats->nonfinite, (long long)h->steps);
  }
  return GP_OK;
}

int gp_theta_count(const gp_handle *h) { return h ? theta_count(h->N, h->NC) : GP_E_ARG; }

int gp_theta_values(const gp_handle *h, double *theta) {
  if (!h || !theta) return GP_E_ARG;
  const gp_params *p = &h->p;
  const int N = h->N, NC = h->NC, L = N - 1;
  int t = 0;
  for (int a = 0; a < N; a++)
    for (int b = a + 1; b < N; b++) {
      theta[t++] = p->dab[a * GP_MAXP + b];
      theta[t++] = p->tau_ab[a * GP_MAXP + b];
      theta[t++] = p->gamma[a * GP_MAXP + b];
    }
  for (int k = 0; k < NC; k++) theta[t++] = p->diffusivity[(L * GP_MAXNC + k) * GP_MAXNC + k];
  theta[t++] = p->tau;
  return GP_OK;
}

int gp_adjoint_forward(gp_handle *h, int64_t nsteps, int64_t every) {
  if (!h || nsteps < 0) return h ? gp_fail(h, GP_E_ARG, "bad nsteps") : GP_E_ARG;
  if (h->p.tempgrady || h->p.shift)
    return gp_fail(h, GP_E_ARG, "the adjoint covers isothermal runs in a fixed frame");
  if (!h->use_march) return gp_fail(h, GP_E_ARG, "the adjoint records single-pass steps (GP_PATH=twopass is set)");
  if (h->fs > 0x7fffffff) return gp_fail(h, GP_E_ARG, "the adjoint indexes cells with 32 bits: grid too large");
  delete h->adj;
  h->adj = new Adjoint();
  Adjoint *A = h->adj;
  if (every <= 0) every = (int64_t)ceil(sqrt((double)(nsteps > 0 ? nsteps : 1)));
  A->nsteps = nsteps;
  A->every = every;
  const int64_t nck = (nsteps + every - 1) / every;
  for (int64_t k = 0; k < nck; k++) {
    double *b = nullptr;
    if (hipMalloc((void **)&b, state_bytes(h)) != hipSuccess)
      return gp_fail(h, GP_E_OOM, "adjoint: checkpoint %lld does not fit in device memory", (long long)k);
    A->ckpt.push_back(b);
    int rc = state_save(h, b);
    if (!rc) rc = gp_step(h, std::min(every, nsteps - k * every), nullptr);
    if (rc) return rc;
  }
  return GP_OK;
}

int gp_adjoint_backward(gp_handle *h, const double *gphi_T, const double *gmu_T, double *gphi_0, double *gmu_0,
                        double *gtheta) {
  if (!h || !gphi_T || !gmu_T || !gphi_0 || !gmu_0 || !gtheta) return h ? gp_fail(h, GP_E_ARG, "null argument") : GP_E_ARG;
  Adjoint *A = h->adj;
  if (!A) return gp_fail(h, GP_E_ARG, "gp_adjoint_forward must run first");
  const int nth = theta_count(h->N, h->NC);
  const size_t se = state_elems(h);
  if (!A->g[0]) {
    bool ok = hipMalloc((void **)&A->g[0], se * sizeof(double)) == hipSuccess &&
              hipMalloc((void **)&A->g[1], se * sizeof(double)) == hipSuccess &&
              hipMalloc((void **)&A->gdphi, (size_t)h->N * h->fs * sizeof(double)) == hipSuccess &&
              hipMalloc((void **)&A->gth, sizeof(double) * nth) == hipSuccess &&
              hipMalloc((void **)&A->cbar, (size_t)h->NC * h->fs * sizeof(double)) == hipSuccess &&
              hipMalloc((void **)&A->qlist, (size_t)h->fs * sizeof(int32_t)) == hipSuccess &&
              hipMalloc((void **)&A->plist, (size_t)h->fs * sizeof(int32_t)) == hipSuccess &&
              hipMalloc((void **)&A->counts, 4 * sizeof(int)) == hipSuccess &&
              hipMemsetAsync(A->counts, 0, 4 * sizeof(int), h->stream) == hipSuccess &&
              hipMalloc((void **)&A->pcode, (size_t)h->fs) == hipSuccess &&
              hipMalloc((void **)&A->bq, (size_t)h->fs) == hipSuccess &&
              hipMalloc((void **)&A->cand, (size_t)h->fs) == hipSuccess;
    for (int64_t j = 0; ok && j < std::min(A->every, A->nsteps); j++) {
      double *b = nullptr;
      ok = hipMalloc((void **)&b, state_bytes(h)) == hipSuccess;
      if (ok) A->seg.push_back(b);
    }
    if (ok && h->precision != GP_FP64) {
      Model<double> m;
      make_model(&h->p, &m);
      ok = hipMalloc((void **)&A->st64, se * sizeof(double)) == hipSuccess &&
           hipMalloc((void **)&A->dphi64, (size_t)h->N * h->fs * sizeof(double)) == hipSuccess &&
           hipMalloc(&A->model64, sizeof m) == hipSuccess &&
           hipMemcpyAsync(A->model64, &m, sizeof m, hipMemcpyHostToDevice, h->stream) == hipSuccess &&
           hipStreamSynchronize(h->stream) == hipSuccess;
    }
    if (!ok) return gp_fail(h, GP_E_OOM, "adjoint: buffers do not fit in device memory");
  }
  int rc;
  HIPCHK(h, hipMemsetAsync(A->g[1], 0, se * sizeof(double), h->stream));
  if ((rc = planes_to_device(h, A->g[1], gphi_T, h->N))) return rc;
  if ((rc = planes_to_device(h, A->g[1] + (size_t)h->N * h->fs, gmu_T, h->NC))) return rc;
  HIPCHK(h, hipMemsetAsync(A->gth, 0, sizeof(double) * nth, h->stream));
  step_fn f = h->ops->step;
  adj_fn fa = h->ops->adj_step;
  widen_fn fw = h->ops->widen;
  const bool wide = h->precision == GP_FP64;
  /* GP_ADJOINT_CHECK=<n>: scan the adjoint after every transposed step (one synchronisation each) and print its
   * largest magnitude every n steps */
  const char *ck = getenv("GP_ADJOINT_CHECK");
  const bool check = ck && *ck && strcmp(ck, "0") != 0;
  const int64_t check_every = check ? atoll(ck) : 0;
  for (int64_t k = (int64_t)A->ckpt.size() - 1; k >= 0; k--) {
    const int64_t n = std::min(A->every, A->nsteps - k * A->every);
    if ((rc = state_load(h, A->ckpt[k]))) return rc;
    for (int64_t j = 0; j < n; j++) {
      if ((rc = state_save(h, A->seg[j]))) return rc;
      if (j + 1 < n && (rc = gp_step(h, 1, nullptr))) return rc;
    }
    for (int64_t j = n - 1; j >= 0; j--) {
      ProfScope ps(h, GP_KERNEL_ADJ);
      if ((rc = state_load(h, A->seg[j]))) return rc;
      if (wide) {
        const double *p64 = (const double *)h->phi[h->cur], *m64 = (const double *)h->mu[h->cur];
        if ((rc = h->ops->adj_dphi(h, p64, m64, (double *)h->dphi, h->model))) return rc;
        rc = fa(h, p64, m64, (const double *)h->dphi, h->model);
      } else {
        if ((rc = fw(h))) return rc;
        if ((rc = h->ops->adj_dphi(h, A->st64, A->st64 + (size_t)h->N * h->fs, A->dphi64, A->model64))) return rc;
        rc = fa(h, A->st64, A->st64 + (size_t)h->N * h->fs, A->dphi64, A->model64);
      }
      if (rc) return rc;
      if (check) {
        const double *p64 = wide ? (const double *)h->phi[h->cur] : A->st64;
        const double *m64 = wide ? (const double *)h->mu[h->cur] : A->st64 + (size_t)h->N * h->fs;
        if ((rc = adj_check(h, k * A->every + j, check_every, p64, m64))) return rc;
      }
    }
  }
  if ((rc = adj_fold(h))) return rc;
  if ((rc = planes_to_host(h, gphi_0, A->g[1], h->N))) return rc;
  if ((rc = planes_to_host(h, gmu_0, A->g[1] + (size_t)h->N * h->fs, h->NC))) return rc;
  HIPCHK(h, hipMemcpyAsync(gtheta, A->gth, sizeof(double) * nth, hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  h->steps = 0; /* the handle holds the initial state again */
  return GP_OK;
}

int gp_profile_enable(gp_handle *h, int enable) {
  if (!h) return GP_E_ARG;
  h->profiling = enable == GP_PROFILE_CALLS ? GP_PROFILE_CALLS : enable ? GP_PROFILE_LAUNCHES : GP_PROFILE_OFF;
  return GP_OK;
}

const char *gp_kernel_name(int kind) {
  static const char *names[GP_NKERNELS] = {"k_phi", "k_mu", "k_stats", "k_march", "k_march_stats", "k_bulk",
                                           "k_dphi", "adj_step", "k_cbulk", "k_ilist", "call", "k_halo"};
  return (kind >= 0 && kind < GP_NKERNELS) ? names[kind] : "?";
}

int gp_profile_read(gp_handle *h, double *avg_ms, int64_t *count) {
  if (!h || !avg_ms || !count) return h ? gp_fail(h, GP_E_ARG, "null argument") : GP_E_ARG;
  HIPCHK(h, hipStreamSynchronize(h->stream));
  double tot[GP_NKERNELS] = {0};
  int64_t cnt[GP_NKERNELS] = {0};
  for (auto &r : h->prof) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, r.second.first, r.second.second) == hipSuccess && r.first >= 0 && r.first < GP_NKERNELS) {
      tot[r.first] += ms;
      cnt[r.first]++;
    }
    (void)hipEventDestroy(r.second.first);
    (void)hipEventDestroy(r.second.second);
  }
  h->prof.clear();
  for (int k = 0; k < GP_NKERNELS; k++) {
    avg_ms[k] = cnt[k] ? tot[k] / (double)cnt[k] : 0.0;
    count[k] = cnt[k];
  }
  return GP_OK;
}

int gp_synchronize(gp_handle *h) {
  if (!h) return GP_E_ARG;
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return GP_OK;
}

}  // extern "C"
