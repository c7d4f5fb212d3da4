/*
 * tomograpy_amd -- C ABI of the MI355X (gfx950) Siddon conic projector /
 * backprojector library.
 *
 * Every entry point is a plain C function over POD structs, raw device
 * pointers and sizes.  The caller owns every buffer (the Python layer uses the
 * torch allocator); the library allocates nothing on the device.  Calls are
 * asynchronous on the caller's HIP
 * stream and return 0 on success or a negative TG_E* code; tg_last_error()
 * returns a thread-local message for the last failing call on this thread.
 *
 * Reference interfaces replaced (tomograpy/C_siddon.c.template = "C:"):
 *   16 generated modules _C_siddon_<ctype>_<obstacle>_<pj> (parse_templates.py:11-26,
 *   setup.py:13-22), each exporting the 7 functions of C:163-172.  Here the
 *   ctype, obstacle and pj kind are fields of tg_call instead of module names.
 */
#ifndef TOMOGRAPY_AMD_H
#define TOMOGRAPY_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Element types (parse_templates.py:11 "float" / "double"). */
#define TG_F32 0
#define TG_F64 1

/* Update kinds (C:915-944: pj, bpj, pjt, bpjt). */
#define TG_PJ 0
#define TG_BPJ 1
#define TG_PJT 2
#define TG_BPJT 3

/* tg_call.flags */
#define TG_OBSTACLE_SUN 1u   /* stop a ray at the unit sphere (C:956-961) */
#define TG_DETERMINISTIC 2u  /* bpj/bpjt: sum each voxel in the reference's serial
                                order (view, row, column), bit-exact to nthread=1 */
#define TG_ZERO_OUT 4u       /* pj/pjt: write every pixel of the launched views as if the
                                image stack had been zeroed first (lo_wrapper.py:26) */
#define TG_ACC_F64 8u        /* bpj/bpjt (atomic, plan and batched): sum the per-step products
                                d * data (rounded in the data type, as C:923-928) in the float64
                                buffer tg_call.acc, then add each voxel's sum, rounded once, to
                                the cube.  The result no longer depends on the order in which
                                rays are summed (up to rare one-ulp rounding ties).  Needs a
                                dense cube: strides (n1*n2, n2, 1) and n0*n1*n2 per time bin. */

/* Error codes */
#define TG_OK 0
#define TG_EINVAL -1     /* bad argument (shape, stride, dtype, kind) */
#define TG_EHIP -2       /* a HIP runtime call failed */
#define TG_EWORKSPACE -3 /* deterministic mode: workspace missing or too small */
#define TG_EGEOMETRY -4  /* a ray's traversal state is NaN (the reference would spin forever) */

/* Per-view header record: 18 doubles, the keys C:1077-1101 reads, in order. */
#define TG_HDR 18
enum {
  TG_H_CRPIX1 = 0, TG_H_CRPIX2, TG_H_CDELT1, TG_H_CDELT2, TG_H_CRVAL1, TG_H_CRVAL2,
  TG_H_R11, TG_H_R12, TG_H_R13, TG_H_R21, TG_H_R22, TG_H_R23, TG_H_R31, TG_H_R32, TG_H_R33,
  TG_H_M1, TG_H_M2, TG_H_M3
};

/* Map cube: the keys C:1058-1075 reads (CDELT, PSHAPE, MMIN, MMAX per axis),
 * the cube shape and its element strides.  stride[3] is the time-bin stride of a
 * 4-D map (C:32-35); nt is the number of time bins (1 for a 3-D map). */
typedef struct tg_map {
  double cdelt[3], pshape[3], mmin[3], mmax[3];
  int32_t n[3];
  int32_t nt;
  int64_t stride[4];
} tg_map;

/* One projector / backprojector invocation (C:460-498). */
typedef struct tg_call {
  int32_t dtype;  /* TG_F32 | TG_F64 */
  int32_t kind;   /* TG_PJ .. TG_BPJT */
  uint32_t flags; /* TG_OBSTACLE_SUN | TG_DETERMINISTIC */
  int32_t tile_j; /* pixels per block row (power of two, 1..256); 0 = n2 rounded up to a
                     power of two, at most 256 */
  tg_map map;
  /* image stack, shape (n1, n2, nv) as in the reference (images on the last axis) */
  int32_t n1, n2, nv, pad0;
  int64_t data_stride[3];
  const void *views; /* device: packed view descriptors, see tg_views_pack */
  void *cube;        /* device: map cube (in for pj, accumulated for bpj) */
  void *data;        /* device: images (accumulated for pj, in for bpj) */
  /* optional operands (NULL = absent) */
  const uint8_t *pixmask;   /* device: nonzero = skip this pixel's ray (C:520) */
  int64_t mask_stride[3];
  const int32_t *tindex;    /* device: time bin per view for pjt/bpjt; NULL = view index (C:935) */
  const void *pix_weight;   /* device: per-pixel weight, data strides (diag O of O.P.M) */
  const void *vox_weight;   /* device: per-voxel weight, vox_stride (diag M of O.P.M) */
  const uint8_t *voxmask;   /* device: nonzero = voxel excluded (P.Mo^T) */
  int64_t vox_stride[3];
  int32_t *status;          /* device int32 or NULL: set to TG_EGEOMETRY if a ray stalls */
  void *workspace;          /* device scratch for TG_DETERMINISTIC */
  size_t workspace_bytes;
  double *acc;              /* device, TG_ACC_F64: n0*n1*n2*nt doubles, zero on entry and
                               zero again when the call's work completes */
  const uint32_t *ray_steps; /* device or NULL, pj/pjt: the Siddon step count of every ray,
                                indexed (t * n1 + i) * n2 + j, as tg_count_steps writes it for
                                this geometry (mask and obstacle included); lets the projector
                                run each ray's loop for a known count (ignored for stacks of
                                2^32 rays or more) */
} tg_call;

/* Size in bytes of the packed view descriptors for nv views of n1 x n2 pixels. */
size_t tg_views_bytes(int32_t dtype, int32_t nv, int32_t n1, int32_t n2);

/* Packs nv header records (nv * TG_HDR doubles) into host memory `out`
 * (tg_views_bytes bytes): the values rounded to the element type as
 * C:1051-1056 does, plus per-view tables of cos/sin of every row angle gamma
 * and column angle lambda computed with libm in double precision, as
 * C:1013-1018 and C:1043-1046 compute them.  The caller copies `out` to the
 * device and passes it as tg_call.views.  Replaces the header marshalling of
 * C:477-487 (PyDict_AsImageHeader, C:1077-1101). */
int tg_views_pack(int32_t dtype, int32_t nv, int32_t n1, int32_t n2, const double *hdr, void *out,
                  size_t out_bytes);

/* All views: replaces conic_full_projector (C:182-205, C:460-498) for the
 * kind/dtype/obstacle selected in *c (siddon.py:85-219 select the module). */
int tg_conic_full_projector(const tg_call *c, void *stream);

/* One view t: replaces conic_image_projector (C:207-248, C:500-528). */
int tg_conic_image_projector(const tg_call *c, int32_t t, void *stream);

/* One ray (pixel i, j of view t): replaces ray_projector (C:250-307).  A pixel
 * whose mask is set is skipped (C:284-286 tests mask == 1; here nonzero). */
int tg_ray_projector(const tg_call *c, int32_t i, int32_t j, int32_t t, void *stream);

/* Unit vectors u[i, j, t, 0:3] (C-contiguous, element type of c->dtype) of
 * views [t0, t1); pixels whose data value is NaN are left untouched (C:638).
 * Replaces full_unit_vector (C:310-330, 582-616) with t0 = 0, t1 = nv and
 * image_unit_vector (C:332-365, 618-648) with t1 = t0 + 1. */
int tg_unit_vector(const tg_call *c, int32_t t0, int32_t t1, void *u, void *stream);

/* Intersection parameters a1/an (C:723-735) from u, views [t0, t1), same
 * layout as u.  Replaces full_intersection_parameters (C:368-399, 650-692)
 * and image_intersection_parameters (C:402-455, 694-720). */
int tg_intersection_parameters(const tg_call *c, int32_t t0, int32_t t1, const void *u, void *a1,
                               void *an, void *stream);

/* Siddon step count (the number of voxel() calls C:575-578 would make) per ray,
 * written to per_ray[t][i][j] (may be NULL), and added to *total (device
 * uint64, may be NULL).  No reference counterpart: used for throughput units. */
int tg_count_steps(const tg_call *c, uint32_t *per_ray, unsigned long long *total, void *stream);

/* Minimum workspace for TG_DETERMINISTIC backprojection of *c: one pixel row
 * of rays per pass.  Larger workspaces take more rows (views) per pass. */
size_t tg_det_workspace_min_bytes(const tg_call *c);

/* Workspace with which a TG_DETERMINISTIC backprojection of all views of *c
 * builds a canonical brick plan inside it (each brick's records in ray order)
 * and sums every brick's voxels in the reference's order through it, instead
 * of the chunked passes; 0 when the geometry does not fit one (more than 2048
 * bricks x time bins of 32 x 16 x 16 voxels).  An upper bound: the call checks
 * the exact need after counting and takes the chunked path if it does not fit.
 * No reference counterpart (the reference's order is its nthread = 1 run). */
size_t tg_det_plan_workspace_bytes(const tg_call *c);

/* Elementwise x[k] *= w[k] over a strided 3-D (or 4-D with nt bins) cube: the
 * voxel-weight half of the Thomson operator's adjoint (models.py:223-235). */
int tg_scale_cube(const tg_call *c, void *stream);

/* c->cube[v] = vox_weight[v] * src[v], or 0 where voxmask[v] != 0 (either
 * operand may be NULL), over the call's cube shape and every time bin of the
 * time kinds; src has its own element strides (src_stride[3] steps the bins)
 * and must not be c->cube.  This is the diag(f) of the Thomson operator
 * (models.py:223-235) and the zero-filled masked unknowns of mask_object
 * (models.py:155-165) applied to the map before a forward projection, so that
 * the projector then reads one value per Siddon step. */
int tg_vox_apply(const tg_call *c, const void *src, const int64_t src_stride[4], void *stream);

/* ---- brick-privatised backprojection (no reference counterpart; replaces the
 * atomic scatter of C:923-928 / C:939-944 for large cubes) ----
 * A plan, built once per geometry (views, map, mask, obstacle, time bins),
 * records each ray's traversal state at every brick it enters.  A
 * backprojection then runs one workgroup per chunk of a brick's records,
 * accumulating the brick in LDS and flushing it with one coalesced atomic per
 * touched voxel.  Lengths and order of operations per step are those of the
 * continuous traversal; only the summation order into a voxel differs (as
 * in the atomic mode). */
typedef struct tg_bpj_plan {
  int32_t brick[3];      /* brick edge per axis, powers of two; 0,0,0 = default: 32 x 16 x 32
                            (132 KiB of LDS), or 16 x 16 x 32 for cubes of edge <= 96; a
                            brick's float64 accumulators, b0 (b1 (b2 + 1) + 1) of them,
                            must fit 160 KiB */
  int32_t chunk_records; /* records per workgroup chunk; 0 = chosen by tg_bpj_plan_count
                            (about 1024 chunks of 2K..64K records) and written back */
  int32_t nb_axis[3];    /* bricks per axis (set by tg_bpj_plan_init) */
  int32_t nbricks;       /* bricks x time bins (set by tg_bpj_plan_init) */
  int32_t order;         /* 0 = records of a brick in trace order, 1 = permuted (spreads the
                            rays of a wavefront over the brick), 2 = trace order with each
                            pass of the backprojection workgroup (1024 records) sorted by
                            decreasing step count (even walks within a wavefront),
                            3 = canonical: ray order (view, row, column), for the
                            deterministic backprojection */
  int32_t trace_tile;    /* pixels per row of the ray tiles the plan build traces (power of
                            two); sets which rays are neighbours in a brick's record list,
                            hence in a wavefront of the backprojection; 0 = 256 */
  int32_t tie_spread;    /* order 2: records with equal step counts keep the order of their
                            pass positions with the low tie_spread bits reversed (0..10;
                            0 = trace order).  Spreading them puts rays that are not
                            neighbours in one wavefront, which pays when a voxel spans
                            several pixels: adjacent rays then add into the same LDS cells */
  int32_t pairs_ready;   /* set by tg_bpj_plan_count for tg_bpj_plan_build: bit 0 when it left
                            each trace workgroup's per-brick record counts in scratch, bit 1
                            when it wrote every ray's step count to ray_steps_out */
  int64_t records;       /* set by tg_bpj_plan_count */
  int64_t chunks;        /* set by tg_bpj_plan_count */
  void *meta;            /* device, caller-owned, meta_bytes */
  size_t meta_bytes;     /* required size, set by tg_bpj_plan_init */
  void *rec;             /* device, caller-owned, rec_bytes */
  size_t rec_bytes;      /* required size, set by tg_bpj_plan_count */
  void *scratch;         /* device, caller-owned, or NULL: scratch_bytes for tg_bpj_plan_count
                            to hand each trace workgroup's per-brick record counts to
                            tg_bpj_plan_build, which then traces every ray once instead of
                            twice; needed only between the two calls */
  size_t scratch_bytes;  /* size that allows it, set by tg_bpj_plan_init (0: not offered) */
  uint32_t *ray_steps_out; /* device or NULL: tg_bpj_plan_count also writes every ray's Siddon
                              step count there, as tg_count_steps does (for the pj/pjt call
                              with the same views, mask and obstacle); tg_bpj_plan_build then
                              runs each ray for its count instead of testing every step
                              against the map and the obstacle, so the counts must stay as
                              the count pass left them until the build */
} tg_bpj_plan;

/* Fills the derived fields and meta_bytes for the geometry of *c (kind bpj or
 * bpjt; cube edges <= 1024).  Order 3 (canonical) plans keep each brick's
 * records in ray order (view, row, column) for the deterministic
 * backprojection: tg_bpj_plan_backproject with TG_DETERMINISTIC sums each
 * voxel in that order, bit for bit the reference's nthread = 1 run.  They need
 * the scratch (scratch_bytes, set here) through count and build, at most 2048
 * bricks x time bins, bricks of at most 8192 voxels (default 32 x 16 x 16)
 * and rows of 256 traced pixels. */
int tg_bpj_plan_init(const tg_call *c, tg_bpj_plan *p);
/* Counts records per brick into p->meta; synchronises the stream; sets
 * records, chunks and rec_bytes. */
int tg_bpj_plan_count(const tg_call *c, tg_bpj_plan *p, void *stream);
/* Writes the records into p->rec (asynchronous). */
int tg_bpj_plan_build(const tg_call *c, const tg_bpj_plan *p, void *stream);
/* cube += P^T data through the plan (asynchronous). */
int tg_bpj_plan_backproject(const tg_call *c, const tg_bpj_plan *p, void *stream);

/* Bytes of the device queue and of the host staging buffer of a plan batch. */
size_t tg_bpj_plan_batch_bytes(int32_t njobs);
/* tg_bpj_plan_backproject for njobs (call, plan) pairs in one launch over all
 * their chunks (one job table, copied from host_staging to queue on the stream;
 * host_staging must stay untouched until that copy has run).  No reference
 * counterpart: the batched-jobs pattern for the backprojector. */
int tg_bpj_plan_batch(const tg_call *calls, const tg_bpj_plan *plans, int32_t njobs, void *queue,
                      void *host_staging, void *stream);

/* ---- device-resident conjugate gradient (lo.acg consumer, srt.py:225) ---- */

/* out += lam * ((identity ? x : 0) + sum_k D_k^T D_k x) on a C-contiguous
 * array of shape[0..3] (unused trailing dims = 1), with D_k the forward
 * difference along axis k in the numpy.diff convention (output shorter by
 * one), k selected by bit k of axes_mask.  Entries whose voxmask byte is
 * nonzero are left untouched (voxmask may be NULL).  With a per-voxel weight
 * w (element type of dtype, same shape; NULL = none) the prior is the height
 * prior D_k diag(w) and the term sum_k diag(w) D_k^T D_k diag(w) x.
 * Replaces the smoothness priors lo.diff of models.py:85-104 (height_prior:
 * w = r^2 from _radius_map) inside the CG normal equations. */
int tg_regularize(int32_t dtype, const int32_t shape[4], uint32_t axes_mask, double lam,
                  int32_t identity, const void *x, void *out, const uint8_t *voxmask,
                  const void *weight, void *stream);

/* Fused CG vector kernels over contiguous length-n vectors (the lo.acg
 * consumer of the projector pair, srt.py:225).  `scal` is a device array of
 * TG_CG_SCAL doubles:
 *   scal[0] = r.r, scal[1] = p.q, scal[2] = new r.r, scal[3] = alpha,
 *   scal[4] = beta, scal[5..15] free for the caller, scal[16..] partial sums.
 * Reductions run in a fixed order, so results are reproducible.
 * tg_cg_dot:       scal[slot] = sum a*b in double (slot 0..15).
 * tg_cg_update:    alpha = scal[0]/scal[1]; x += alpha p; r -= alpha q;
 *                  scal[2] = r.r.
 * tg_cg_direction: beta = scal[2]/scal[0]; p = r + beta p; scal[0] = scal[2]. */
#define TG_CG_SCAL (16 + 1024)
int tg_cg_dot(int32_t dtype, int64_t n, const void *a, const void *b, double *scal, int32_t slot,
              void *stream);
int tg_cg_update(int32_t dtype, int64_t n, void *x, void *r, const void *p, const void *q,
                 double *scal, void *stream);
int tg_cg_direction(int32_t dtype, int64_t n, const void *r, void *p, double *scal, void *stream);

/* The smoothness priors of the regularised models (lo.diff, models.py:85-104,
 * and the temporal term of stsrt, models.py:140) in differentiable form, over
 * a C-contiguous array of shape[4] (trailing 1s for fewer axes):
 *   scal[slot] = sum_k weights[k] * sum_m (u[m + e_k] - u[m])^2
 * (the numpy.diff convention; squares and sums in double, in a fixed order)
 * and, when grad is not NULL,
 *   grad = 2 diag(v) sum_k weights[k] D_k^T D_k u,
 * where u = v x with the per-voxel weight v = voxel_weight (the height prior
 * r^2 of models.py:90-93; same shape and element type as x), or u = x and
 * v = 1 when voxel_weight is NULL.  Axes with weights[k] == 0 are skipped;
 * scal as for tg_cg_dot. */
int tg_smoothness(int32_t dtype, const int32_t shape[4], const double weights[4], const void *x,
                  const void *voxel_weight, void *grad, double *scal, int32_t slot,
                  void *stream);

/* ---- segmented CG: the vector work of many concurrent solves in one launch
 * (no reference counterpart; used by the reconstruction server) ---- */

/* One solve's vectors (contiguous, n elements each, four distinct arrays that do not overlap)
 * and its regulariser. */
typedef struct tg_cg_seg {
  void *x, *r, *p, *q;
  const uint8_t *voxmask; /* may be NULL */
  const void *weight;     /* per-voxel weight of the height prior (see tg_regularize), or NULL */
  double *scal;           /* this solve's CG scalars, layout of tg_cg_update */
  int64_t n;
  int64_t block0;         /* first launch block of this segment: sum of tg_cg_seg_blocks(n) before it */
  int32_t shape[4];       /* regulariser array shape (C-contiguous, trailing 1s) */
  uint32_t axes;          /* regulariser difference axes (bit k = axis k) */
  int32_t identity;       /* add the identity term */
  double lam;
} tg_cg_seg;

#define TG_CG_OP_REGULARIZE 0 /* q += lam ((identity ? p : 0) + sum_k D_k^T D_k p), weighted as
                                 tg_regularize when tg_cg_seg.weight is set */
#define TG_CG_OP_DOT_PQ 1     /* scal[1] = p.q */
#define TG_CG_OP_UPDATE 2     /* alpha = scal[0]/scal[1]; x += alpha p; r -= alpha q; scal[2] = r.r */
#define TG_CG_OP_DIRECTION 3  /* beta = scal[2]/scal[0]; p = r + beta p; scal[0] = scal[2] */
#define TG_CG_OP_INIT 4       /* scal[0] = scal[5] = r.r */
#define TG_CG_OP_ZERO_Q 5     /* q = 0 */
#define TG_CG_OP_REGULARIZE_DOT 6 /* REGULARIZE, then DOT_PQ, in one pass over the elements */

/* Launch blocks a segment of n elements takes. */
int64_t tg_cg_seg_blocks(int64_t n);

/* Applies `op` to every segment of the device table segs[0..nseg) in one
 * launch of nblocks blocks (the sum of the segments' tg_cg_seg_blocks);
 * `partials` is device scratch of nblocks doubles.  Reductions run in a fixed
 * order per segment. */
int tg_cg_batch(int32_t dtype, int32_t op, const tg_cg_seg *segs, int32_t nseg, int64_t nblocks,
                double *partials, void *stream);

/* ---- batched uneven jobs in one persistent launch (no reference counterpart:
 * the reference runs one conic_full_projector call per job) ---- */

/* Schedule of a batched launch. */
typedef struct tg_batch_plan {
  int32_t tile_j;   /* pixels per workgroup row for every job (power of two); 0 = each job's */
  int32_t blocks;   /* persistent workgroups; 0 = 8 per compute unit */
  int32_t schedule; /* 0 = jobs in the given order, 1 = longest rays first */
  int32_t pad;
} tg_batch_plan;

/* Bytes of the device queue buffer (and of the host staging buffer) for njobs. */
size_t tg_batch_queue_bytes(int32_t njobs);

/* Runs calls[0..njobs) -- all TG_PJ/TG_PJT or all TG_BPJ/TG_BPJT, one dtype,
 * per-job geometry, sizes, obstacle and operands -- as one persistent launch.
 * The job table is written to `host_staging` (tg_batch_queue_bytes bytes,
 * caller-owned; keep it unchanged until the stream has passed this call) and
 * copied to `queue` (device, same size, caller-owned) on `stream`.  Each job's
 * result equals its one-job call's (the forward projector bit for bit). */
int tg_batch_run(const tg_call *calls, int32_t njobs, const tg_batch_plan *plan, void *queue,
                 void *host_staging, void *stream);

/* Host-only: writes the job table tg_batch_run would upload into host_staging
 * and launches nothing (for inspection and tests).  Header words: [1] = total
 * tiles, [2] = njobs; jobs follow at byte 256 in queue order. */
int tg_batch_plan_table(const tg_call *calls, int32_t njobs, const tg_batch_plan *plan,
                        void *host_staging);

/* Layout of one job-table entry for `dtype`: its size, the byte offset of its
 * first-tile index (uint64) and of its cube pointer. */
int tg_batch_job_layout(int32_t dtype, size_t *job_bytes, size_t *tile_begin_offset,
                        size_t *cube_offset);

/* ---- an entry point that allocates nothing ----
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): tg_conic_full_projector for bpj/bpjt with the float64
 * accumulation of TG_ACC_F64, whose accumulator is a device array inside the
 * library (tg_static_acc_capacity() doubles per GPU, held until the process
 * ends) instead of c->acc, which is ignored.  The cube must be dense (as
 * TG_ACC_F64 requires) and n0*n1*n2 (times nt for bpjt) at most the capacity,
 * else TG_EINVAL; TG_DETERMINISTIC is rejected.  pj/pjt calls run as in
 * tg_conic_full_projector.  The call clears the part of the array it uses
 * with a kernel, then sums into it; every call on one GPU shares that GPU's
 * array, so calls that overlap on one GPU (other streams, other threads) race
 * and must be serialised by the caller. */
int tg_conic_full_projector_static(const tg_call *c, void *stream);
/* Doubles in the static accumulator (2^24). */
int64_t tg_static_acc_capacity(void);

/* Thread-local message describing the last error on this thread. */
const char *tg_last_error(void);

/* Library version string. */
const char *tg_version(void);

#ifdef __cplusplus
}
#endif

#endif /* TOMOGRAPY_AMD_H */
