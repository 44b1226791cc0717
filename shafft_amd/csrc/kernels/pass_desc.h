/* Descriptor of one LDS-resident FFT pass and the launch ABI between the
 * host plan code (C++) and the HIP translation units.  Plain C so that both
 * sides, and any future caller, see the same layout.
 *
 * A pass transforms `batch` lines of length `n`.  Line b is split into three
 * mixed-radix digits (b0, b1, b2) with b2 = b % n2 fastest and
 * b1 = (b / n2) % n1; element k of line b is read at
 *     in [b0*is0 + b1*is1 + b2*is2 + k*iks]
 * and written at
 *     out[b0*os0 + b1*os1 + b2*os2 + k*oks]      (strides in complex elements).
 * Independent in/out strides are what fuse the slab transposes of a
 * multi-pass transform into the loads and stores of the adjacent pass.
 */
#ifndef SHAFFT_AMD_PASS_DESC_H
#define SHAFFT_AMD_PASS_DESC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SA_MAX_STAGES 16

typedef struct {
  uint32_t d; /* divisor */
  uint32_t m; /* magic multiplier */
  uint32_t s; /* post-shift */
} sa_fastdiv_t;

typedef struct {
  const void* in;
  void* out;
  int64_t batch;
  int64_t n1, n2;
  int64_t is0, is1, is2, iks;
  int64_t os0, os1, os2, oks;

  int32_t n;       /* transform length of this pass                    */
  int32_t nstages; /* Stockham stages (0 => plain strided copy)        */
  int32_t radix[SA_MAX_STAGES];
  int32_t ns[SA_MAX_STAGES];    /* product of the radices before stage s   */
  int32_t twoff[SA_MAX_STAGES]; /* offset (elements) into `tw` for stage s */
  sa_fastdiv_t nsdiv[SA_MAX_STAGES];
  sa_fastdiv_t nbdiv[SA_MAX_STAGES]; /* n / radix[s]: butterflies per line */
  const void* tw; /* per-stage forward twiddles, same precision as data   */

  int32_t C;         /* lines per workgroup                              */
  int32_t T;         /* threads per line                                 */
  int32_t emax;      /* generic kernel: elements per thread (16, or 8 for long Bluestein rows) */
  int32_t ld_kfast;  /* 1: lanes walk k on load, 0: lanes walk lines     */
  int32_t st_kfast;  /* same for the store                               */
  int32_t ldsstride; /* complex elements per line slot in LDS (padded)   */
  int32_t reg_first; /* stage 0 straight from the loaded registers       */
  int32_t reg_last;  /* last stage straight into the stored registers    */
  int32_t nt;        /* non-temporal global loads/stores of the data     */
  int32_t logC;      /* log2(C) for the fixed-length kernel (C a power of 2) */
  int32_t tw_lds;    /* fixed kernel: stage-twiddle elements staged in LDS (0: global) */
  int32_t xcd_order; /* fixed kernel: 1 = each XCD walks its own contiguous tile range; >= 2 = rows of that many tiles walked down the rows first */
  int32_t lds_pad;   /* generic kernel: 1 = LDS rows padded k -> k + k/16, 0 = unpadded */
  sa_fastdiv_t Tdiv, Cdiv, n1div, n2div;

  /* Optional post-twiddle on the store: out *= W_N^{+-(digit * k) mod N}
   * with digit = (b0, b1, b2)[ptw_level]; W from a two-level table
   * W^e = hi[e >> ptw_shift] * lo[e & ((1 << ptw_shift) - 1)].           */
  int32_t ptw_level; /* -1: none */
  int32_t ptw_conj;  /* 1: multiply by conj(W) instead                   */
  int32_t ptw_shift;
  int64_t ptw_N;
  const void* ptw_lo;
  const void* ptw_hi;

  /* Element-index hooks.  For element k of line b with digits (b0,b1,b2):
   *   m_ld = k * ld_kmul + digit[ld_level] * ld_dmul + digit[ld_level2] * ld_dmul2
   * (a term is absent when its level is < 0)
   * the element is read only if m_ld < ld_limit (else taken as zero) and,
   * when `pre` is set, multiplied by pre[m_ld].  Likewise on the store with
   * st_*: written only if m_st < st_limit, multiplied by post[m_st].
   * Plain passes use kmul = 1, level = -1, limit = n.                     */
  int64_t ld_kmul, ld_dmul, ld_dmul2, ld_limit;
  int64_t st_kmul, st_dmul, st_dmul2, st_limit;
  int32_t ld_level, ld_level2, st_level, st_level2;
  const void* pre;
  const void* post;
  /* Bluestein mode: the pass runs forward stages, multiplies by blu_h[k]
   * (FFT of the conjugate chirp, already divided by n), then runs backward
   * stages, all inside LDS.  The direction is carried by pre/post.       */
  const void* blu_h;
  /* Per-line filter: line b multiplies by blu_h + digit[blu_level] *
   * blu_hstride (blu_level < 0: the same filter for every line).         */
  int32_t blu_level;
  int64_t blu_hstride;
} sa_pass_t;

/* Launch one pass.  prec: 0 = complex float, 1 = complex double.
 * sign: -1 forward, +1 backward.  Returns a hipError_t value. */
int sa_launch_pass(const sa_pass_t* p, int prec, int sign, void* stream);
/* Workgroups of the pass kernel per CU for p's launch shape, with (twl = 1)
 * or without its stage twiddles staged in LDS; 0 if it cannot launch. */
int sa_pass_occupancy(const sa_pass_t* p, int prec, int twl);

/* Fixed-length persistent kernel (power-of-two n in [32, 8192], no load or
 * store hooks other than the four-step twiddle).  `grid` is the
 * number of persistent workgroups. */
int sa_fixed_supported(int n);
int sa_fixed_elems(int n, int prec); /* elements per thread (n / threads per line) */
int sa_fixed_max_threads(int prec); /* workgroup size bound */
int sa_launch_fixed(const sa_pass_t* p, int prec, int sign, int grid, void* stream);
/* Resident workgroups per CU of that kernel for a block size and LDS size. */
int sa_fixed_occupancy(int n, int prec, int sign, int nthr, size_t lds);

/* Three-pass Bluestein kernels on compile-time power-of-two lengths
 * (fft_blue.hip): the chirped, zero-padded column pass, the row pass
 * (forward DFT, per-row filter, inverse DFT, conjugate four-step twiddle)
 * and the chirped, truncating inverse column pass.  Same descriptor
 * semantics as the generic pass; tiles of C = 2^logC lines, one tile per
 * workgroup, T = n / sa_blue_elems(...) threads per line (p->emax). */
enum { SA_BLUE_COLS_IN = 1, SA_BLUE_ROWS = 2, SA_BLUE_COLS_OUT = 3 };
int sa_blue_supported(int n, int kind);
/* Elements per thread of that kernel for n-point lines in tiles of `lines`. */
int sa_blue_elems(int n, int prec, int kind, int lines);
int sa_launch_blue(const sa_pass_t* p, int prec, int kind, void* stream);

/* Batched persistent launch: many passes of many transforms in one grid.
 * Workgroups pull items (pass, tile) from a device queue in list order;
 * an item waits until every tile of the pass it depends on has completed.
 * The list must be ordered so that every item's dependency pass has all of
 * its tiles earlier in the list (level-major order does this).          */
typedef struct {
  uint32_t pass;
  uint32_t tile;
} sa_batch_item_t;

typedef struct {
  const sa_pass_t* passes;      /* pass descriptors (device)                  */
  const uint32_t* pass_tiles;   /* tiles per pass (device)                    */
  const int32_t* pass_dep;      /* pass each pass waits for, -1 none (device) */
  const int8_t* pass_sign;      /* -1 / +1 per pass (device)                  */
  uint32_t* done;               /* completed tiles per pass, zero at launch   */
  uint32_t* head;               /* queue head, zero at launch                 */
  uint32_t* err;                /* set to 1 if a dependency wait timed out    */
  uint64_t wait_ticks;          /* wall-clock ticks before a wait gives up    */
  const sa_batch_item_t* items; /* work list (device)                         */
  uint32_t nitems;
  uint32_t npasses;
  /* 1: the list holds one dependency level only (the launch boundary is the
   * synchronisation): no waits and no per-tile release.                   */
  uint32_t independent;
  /* Radix-8/16 stages of complex-float passes: 0 FP32 on the vector ALUs,
   * 1 bfloat16 hi/lo split on the matrix cores, 2 float16 hi/lo split. */
  int32_t mfma;
  /* 1: `head` and `done` are already zero for this launch (the caller
   * wrote them), so no clearing kernel runs before it.                     */
  int32_t precleared;
  /* Work queues (1 or 8; 8 only with `independent`): the list is split into
   * `queues` contiguous ranges, range q taken by the workgroups of XCD q
   * (workgroup b runs on XCD b % 8) through its own counter head[q], so
   * neighbouring tiles meet in one L2; a workgroup whose range is used up
   * takes items from the other ranges.  `head` holds `queues` counters.    */
  uint32_t queues;
} sa_batch_t;

/* Zero the queue head and the per-pass counters (a kernel, so it can sit in
 * a captured graph; skipped when b->precleared), then run the persistent
 * batch kernel with `grid` workgroups of `block` threads and `lds` bytes of
 * dynamic LDS. */
int sa_launch_batch(const sa_batch_t* b, int prec, int grid, int block, size_t lds, void* stream);
/* Resident batch workgroups per CU for a block size and LDS size. */
int sa_batch_occupancy(int prec, int block, size_t lds);

/* x[i] *= f for i < count complex elements. */
int sa_launch_scale(void* data, double f, uint64_t count, int prec, void* stream);

/* Largest dynamic LDS a pass may use (bytes). */
int sa_max_lds_bytes(void);

/* Device address and size of the allocation-free execute's scratch array on
 * the current device (resolved once per device, then cached). */
int sa_noalloc_scratch(void** ptr, size_t* bytes);

#ifdef __cplusplus
}
#endif

#endif
