/*
 * TEST INFRASTRUCTURE ONLY -- the CPU oracle for the Siddon hot path.
 *
 * A plain C/OpenMP restatement of tomograpy/C_siddon.c.template (the
 * reference's only native component).  It is the checker of the HIP product
 * library and, timed on the host cores, the CPU baseline reported beside the
 * GPU numbers.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it; the product never does.
 *
 * Build: oracle/Makefile (gcc -O2 -fopenmp -ffp-contract=off).  The reference
 * template itself is not built here: it needs its Python-2 code generator
 * (tomograpy/parse_templates.py) and the Python-2 C-API (Py_InitModule).
 *
 * Decomposition as in C:460-498: OpenMP over images, serial pixel loops,
 * serial along-ray steps, atomic updates (C:915-944).
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>

/* C:17 -- "To avoid division by zero", in solar radii. */
#define TGO_INF 100000

enum { TGO_PJ = 0, TGO_BPJ = 1, TGO_PJT = 2, TGO_BPJT = 3 };

/* Layout of one per-view header record (18 doubles): the keys C:1077-1101
 * reads, in that order. */
enum {
  TGO_CRPIX1 = 0, TGO_CRPIX2, TGO_CDELT1, TGO_CDELT2, TGO_CRVAL1, TGO_CRVAL2,
  TGO_R11, TGO_R12, TGO_R13, TGO_R21, TGO_R22, TGO_R23, TGO_R31, TGO_R32, TGO_R33,
  TGO_M1, TGO_M2, TGO_M3, TGO_HDR
};

/* Map description: the keys C:1058-1075 reads plus the array geometry. */
typedef struct {
  double cdelt[3], pshape[3], mmin[3], mmax[3];
  int32_t n[3];
  int32_t pad;
  int64_t s[4]; /* element strides of the cube; s[3] is the time axis */
} tgo_map;

/* Image-stack description: shape (n1, n2, nv), element strides, headers. */
typedef struct {
  int32_t n1, n2, nv, pad;
  int64_t s[3];           /* element strides of data */
  int64_t ms[3];          /* element strides of mask */
  const double *hdr;      /* nv * TGO_HDR */
  const uint8_t *mask;    /* NULL or nonzero = skip pixel (C:520) */
  const int32_t *tindex;  /* NULL or time index per view for pjt/bpjt (C:931-944) */
} tgo_data;

#define CT float
#define SFX f32
#include "siddon_oracle_core.h"
#undef CT
#undef SFX

#define CT double
#define SFX f64
#include "siddon_oracle_core.h"
#undef CT
#undef SFX

int tgo_num_threads(void) { return omp_get_max_threads(); }
