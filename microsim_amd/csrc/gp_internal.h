/*
 * gp_internal.h -- private declarations shared by libgp_hip's translation units: the handle, the adjoint's device
 * buffers, error reporting and the kernel sets. Not a public header.
 *
 * The kernels are templates over the number of phases N and of mu fields NC. Each supported (N, NC) pair is
 * instantiated in a translation unit of its own (gp_ops.hip compiled once per pair, so the build runs them in
 * parallel); gp_hip.hip holds the C ABI and reaches the kernels of a handle through its GpOps table.
 */
#pragma once

#include <hip/hip_runtime.h>

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <utility>
#include <vector>

#include "gp_common.h"
#include "gp_frame.h"
#include "gp_hip.h"
#include "gp_model.h"

/* The supported (phases, mu fields) pairs; the build compiles gp_ops.hip once per pair listed here. */
#define GP_COMBOS(X) X(2, 1) X(3, 2) X(3, 1) X(2, 2) X(4, 2) X(4, 1) X(3, 3)

namespace {

constexpr int kBX = 64;          /* threads along y (one wavefront) */
constexpr int kBY = 4;           /* rows per block */
constexpr int kNStatMax = 3 * GP_MAXP + 3 * GP_MAXNC + 1;
#ifndef GP_LIST_BLOCKS
#define GP_LIST_BLOCKS 2048
#endif
constexpr int kListBlocks = GP_LIST_BLOCKS; /* largest grid of the list-driven kernels (grid-stride loops), 4 wavefronts each */
#ifndef GP_ADJ_BLOCKS
#define GP_ADJ_BLOCKS 1024
#endif
constexpr int kAdjBlocks = GP_ADJ_BLOCKS;   /* grid of the adjoint's reverse-mode list kernels */

}  // namespace

struct GpOps;

struct gp_handle {
  gp_params p;
  int precision;
  int N, NC;
  int64_t rx, ry, ps, fs;   /* rows, row pitch (elements), plane size (elements) */
  hipStream_t stream;
  int device;
  void *phi[2], *mu[2], *dphi;
  void *model;              /* Model<R> in device memory */
  /* GP_FP32C: the second parts of the state of each of the two buffers: the FP32 low parts of phi (N planes,
   * c32_plo) then the 8-bit extensions of mu (NC planes of one byte per element, c32_mlo), at the element offsets
   * of the FP32 planes; the FP64 model the interface cells step with; FP64 copies of the decoded state (allocated
   * on first use: smoothing, delta-phi downloads, invariants). dphi holds FP64 planes in this mode. */
  int8_t *lo[2];
  void *model64;
  double *wide[2];
  /* non-isothermal runs (tempgrady) and the moving frame (shift), gp_frame.h: T of every allocated row after the
   * last step, the row coefficients of the current step's two sweeps (RowThermo<R>), the frame's interface row
   * (device) and the cells the frame has moved since gp_upload */
  double *trow;
  void *thp, *thc;
  int *front;
  int64_t shift_off;
  gpk::ThermoParams tp;
  double *partials;         /* [statistic][nparts_max], statistic-major */
  double *stats_dev;        /* cumulative stats (kNStatMax) */
  double *inv_dev;          /* solute[NC], energy of the current state */
  int cur;
  int64_t nblocks_mu;
  int64_t steps;
  int64_t nparts_max;       /* rows of the partials buffer */
  int march_seg, march_nstrips, march_nsegs; /* tile grid of the single-pass step (see gp_march.h) */
  int64_t march_nwaves;     /* tiles = wavefronts of k_march and of k_bulk */
  int cell_seg, cell_nsegs;  /* tile grid of k_cbulk (same strips as the march, shorter segments) */
  int64_t cell_nwaves;
  int32_t *clist, *ilist;    /* delta-phi candidates and interface cells of the current step (cell offsets) */
  int *lcount;               /* their lengths: two pairs, used by alternate steps (lcount + 2 * lsel) */
  int lsel;
  int lgrid_d, lgrid_i;      /* grids of this handle's delta-phi and interface list kernels (0: not yet set) */
  int *flags[2];            /* per-tile purity flags of the last two steps */
  int8_t *codes[2];         /* phase codes (pure phase b, or -1) of every cell: codes[ccur] of the current state */
  int ccur;
  bool codes_valid;         /* codes match the current state */
  int fcur;                 /* flags[fcur] describes the current state */
  bool flags_valid;         /* flags[fcur] is valid for the current state (else no tile is taken as deep) */
  bool use_march;           /* single-pass steps (default) or the two-pass form */
  bool use_cells;           /* single-pass update by cell class (k_cbulk + k_ilist, default) or by march tiles */
  bool prev_valid;          /* the other buffer holds the state before the last step */
  gpk::Maps maps;
  bool pow2;
  int profiling;            /* 0 off, GP_PROFILE_LAUNCHES: events around every launch, GP_PROFILE_CALLS: around
                               every gp_step / gp_smooth call */
  struct Adjoint *adj;      /* checkpoints and buffers of the adjoint (gp_adjoint_forward / _backward) */
  const struct GpOps *ops;  /* kernels instantiated for this handle's phases, components and precision */
  std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> prof;
  char err[512];
};

/* bracket one launch (kind GP_KERNEL_*, in GP_PROFILE_LAUNCHES mode) or one whole call (kind GP_KERNEL_CALL, in
 * GP_PROFILE_CALLS mode) with events on the handle's stream */
struct ProfScope {
  gp_handle *h;
  hipEvent_t e0, e1;
  int kind;
  ProfScope(gp_handle *hh, int k) : h(hh), e0(nullptr), e1(nullptr), kind(k) {
    const int want = k == GP_KERNEL_CALL ? GP_PROFILE_CALLS : GP_PROFILE_LAUNCHES;
    if (h->profiling == want && hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess &&
        hipEventRecord(e0, h->stream) != hipSuccess)
      release();
  }
  ~ProfScope() {
    if (e1) {
      if (hipEventRecord(e1, h->stream) == hipSuccess) h->prof.push_back({kind, {e0, e1}});
      else release();
    }
  }
  /* a launch whose events could not be recorded is simply not timed */
  void release() {
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    e0 = e1 = nullptr;
  }
};

/* record the message of a failing call on h (or of gp_create when h is NULL) and return code */
int gp_fail(gp_handle *h, int code, const char *fmt, ...);

#define HIPCHK(h, call)                                                                         \
  do {                                                                                          \
    hipError_t e_ = (call);                                                                     \
    if (e_ != hipSuccess) return gp_fail((h), GP_E_HIP, "%s: %s", #call, hipGetErrorString(e_)); \
  } while (0)

enum { MODE_SMOOTH = 0, MODE_TWOPASS = 1, MODE_MARCH = 2, MODE_DPHI_PREV = 3, MODE_MARCH_STATS = 4 };

inline size_t esize(const gp_handle *h) { return h->precision == GP_FP64 ? sizeof(double) : sizeof(float); }
inline size_t state_elems(const gp_handle *h) { return (size_t)(h->N + h->NC) * h->fs; }
/* bytes of the second parts of one GP_FP32C state (phi low parts, mu extensions) */
inline size_t c32_lo_bytes(const gp_handle *h) { return (size_t)h->fs * (h->N * sizeof(float) + h->NC); }
inline float *c32_plo(const gp_handle *h, int b) { return (float *)h->lo[b]; }
inline int8_t *c32_mlo(const gp_handle *h, int b) { return h->lo[b] + (size_t)h->N * h->fs * sizeof(float); }
/* bytes of one stored state: the planes, plus the second parts in GP_FP32C */
inline size_t state_bytes(const gp_handle *h) {
  return state_elems(h) * esize(h) + (h->precision == GP_FP32C ? c32_lo_bytes(h) : 0);
}
/* element size of the delta-phi planes (FP64 in GP_FP32C, where the interface cells step in FP64) */
inline size_t dphi_esize(const gp_handle *h) { return h->precision == GP_FP32 ? sizeof(float) : sizeof(double); }

/* ------------------------------------------------------------------------------------------------ adjoint */

/* Device buffers of the adjoint: checkpoints every `every` steps, the recomputed states of one segment, the
 * state adjoints of two consecutive steps, the delta-phi adjoint, the parameter gradient and the cell lists of the
 * reverse-mode stages. */
struct Adjoint {
  int64_t nsteps = 0, every = 0;
  std::vector<double *> ckpt;
  std::vector<double *> seg;
  double *g[2] = {nullptr, nullptr};
  double *gdphi = nullptr, *gth = nullptr, *cbar = nullptr;
  int32_t *qlist = nullptr, *plist = nullptr; /* cells of the reverse-mode stages Q and P */
  /* FP32 handles: the current state widened to FP64, its delta-phi in FP64 and the FP64 model, at which every
   * transposed step is evaluated (the FP64 model's derivative at the FP32 trajectory) */
  double *st64 = nullptr, *dphi64 = nullptr;
  void *model64 = nullptr;
  int *counts = nullptr;                      /* their lengths: [0] stage Q, [1] and [2] stage P (alternating) */
  int psel = 0;                               /* the P slot the next adj_dphi fills is counts[1 + psel] */
  int pslot = 1;                              /* the P slot the last adj_dphi filled */
  unsigned long long *scan = nullptr;         /* GP_ADJOINT_CHECK: first non-finite index, max |g| bits, theta flag */
  /* per allocated cell of the evaluation state: the pure phase of phi (b when phi = e_b exactly, else -1) and the
   * bulk phase of stage Q (bulk_q, -1 for the other cells) */
  int8_t *pcode = nullptr, *bq = nullptr;
  int8_t *cand = nullptr; /* 1 at the delta-phi candidates (k_adj_codes); every other cell's delta-phi is +0 */
  ~Adjoint() {
    for (double *b : ckpt) (void)hipFree(b);
    for (double *b : seg) (void)hipFree(b);
    void *bufs[] = {g[0], g[1], gdphi, gth, cbar, qlist, plist, counts, st64, dphi64, model64, scan, pcode, bq, cand};
    for (void *b : bufs)
      if (b) (void)hipFree(b);
  }
};

typedef int (*step_fn)(gp_handle *, int);
typedef int (*inv_fn)(gp_handle *);
typedef int (*adj_fn)(gp_handle *, const double *, const double *, const double *, const void *);
typedef int (*adj_dphi_fn)(gp_handle *, const double *, const double *, double *, const void *);
typedef int (*widen_fn)(gp_handle *);
typedef int (*inv_on_fn)(gp_handle *, const void *, const void *, int64_t, int64_t, int64_t, double *, double *,
                         hipStream_t);

/* the kernels of one (N, NC, precision): a step (by mode), the solute/energy invariants (of the handle's state, or
 * of any state with given scratch and stream), one transposed step of the adjoint and the FP32 -> FP64 widening the
 * FP32 adjoint evaluates its Jacobians at */
struct GpOps {
  step_fn step;
  inv_fn invariants;
  adj_fn adj_step;
  widen_fn widen;
  inv_on_fn invariants_on; /* the invariants of a given state, with given partials buffer and stream */
  adj_dphi_fn adj_dphi;    /* the adjoint's delta-phi of an evaluation state and its P candidates */
};

#define GP_DECLARE_OPS(a, b) const GpOps *gp_ops_##a##_##b(int precision);
GP_COMBOS(GP_DECLARE_OPS)
#undef GP_DECLARE_OPS
