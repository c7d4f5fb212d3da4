/*
 * TEST INFRASTRUCTURE ONLY.  This file is part of the CPU oracle: a plain C
 * restatement of TomograPy's Siddon conic projector/backprojector used as the
 * checker of the HIP product path.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it.
 *
 * Included twice by siddon_oracle.c with CT = float / double and SFX = f32 /
 * f64.  Every function below restates one function of
 * tomograpy/C_siddon.c.template (cited as C:<line>) with the same expressions
 * in the same order, so that with -ffp-contract=off the float arithmetic is the
 * reference's arithmetic.
 */

#define CAT_(a, b) a##_##b
#define CAT(a, b) CAT_(a, b)
#define FN(name) CAT(name, SFX)

/* C:1043-1046 -- pixel index to physical angle, in CT, FITS 1-based pixels. */
static inline CT FN(pixel2physical)(unsigned int id, CT s, CT p, CT v)
{
  return (id - s + 1) * p + v;
}

/* C:1013-1032 -- cos/sin are libm double functions even when CT is float;
 * the double products are rounded to CT on assignment. */
static inline void FN(rotated_unit_vector)(CT lambda, CT gamma, const CT R[9], CT u0[3])
{
  CT u2[3];
  u2[0] = cos(lambda) * cos(gamma);
  u2[1] = cos(lambda) * sin(gamma);
  u2[2] = sin(lambda);
  u0[0] = R[0] * u2[0] + R[1] * u2[1] + R[2] * u2[2];
  u0[1] = R[3] * u2[0] + R[4] * u2[1] + R[5] * u2[2];
  u0[2] = R[6] * u2[0] + R[7] * u2[1] + R[8] * u2[2];
}

/* C:969-977 */
static inline CT FN(min3)(CT x, CT y, CT z)
{
  if ((x < y) && (x < z)) return x;
  else if ((y <= x) && (y < z)) return y;
  else return z;
}

/* C:979-987 */
static inline CT FN(max3)(CT x, CT y, CT z)
{
  if ((x > y) && (x > z)) return x;
  else if ((y >= x) && (y > z)) return y;
  else return z;
}

/* C:989-997 */
static inline int FN(sign)(CT x)
{
  if (x > 0) return 1;
  else if (x < 0) return -1;
  else return 0;
}

/* Per-CT map header: C:50-56 with values rounded to CT as C:1051-1075 does. */
typedef struct {
  CT p[3], d[3], min[3], max[3];
} FN(mhdr);

/* Per-CT image header: C:58-65, rounded as C:1077-1101 does. */
typedef struct {
  CT s[2], p[2], v[2], R[9], M[3];
} FN(ihdr);

static void FN(load_map)(const tgo_map *m, FN(mhdr) *mh)
{
  for (int k = 0; k < 3; k++) {
    mh->p[k] = (CT)m->cdelt[k];
    mh->d[k] = (CT)m->pshape[k];
    mh->min[k] = (CT)m->mmin[k];
    mh->max[k] = (CT)m->mmax[k];
  }
}

static void FN(load_image)(const double *h, FN(ihdr) *ih)
{
  ih->s[0] = (CT)h[TGO_CRPIX1]; ih->s[1] = (CT)h[TGO_CRPIX2];
  ih->p[0] = (CT)h[TGO_CDELT1]; ih->p[1] = (CT)h[TGO_CDELT2];
  ih->v[0] = (CT)h[TGO_CRVAL1]; ih->v[1] = (CT)h[TGO_CRVAL2];
  for (int k = 0; k < 9; k++) ih->R[k] = (CT)h[TGO_R11 + k];
  for (int k = 0; k < 3; k++) ih->M[k] = (CT)h[TGO_M1 + k];
}

/* C:723-735 */
static inline void FN(intersection_parameters)(const FN(mhdr) *mh, const CT M[3], const CT u0[3],
                                               CT a1[3], CT an[3])
{
  for (int k = 0; k < 3; k++) {
    if (u0[k] != 0) {
      a1[k] = (mh->min[k] - M[k]) / u0[k];
      an[k] = (mh->max[k] - M[k]) / u0[k];
    } else {
      a1[k] = -TGO_INF;
      an[k] = TGO_INF;
    }
  }
}

/* Ray state after the setup of C:561-573. */
typedef struct {
  CT a1[3], D[3], pabs[3], ac, amin, amax;
  int iv[3], update[3];
  int hit;
} FN(ray);

/* C:561-573 with C:737-809: intersections, steps, update, first voxel,
 * next plane and distances.  Float -> unsigned casts follow x86-64 gcc code
 * generation for (unsigned int)x: truncation to 64-bit then the low 32 bits. */
static inline void FN(ray_setup)(const FN(mhdr) *mh, const unsigned int n[3], const CT M[3],
                                 const CT u0[3], FN(ray) *r)
{
  CT an[3], Imin[3], Imax[3], p[3];
  unsigned int next[3];
  FN(intersection_parameters)(mh, M, u0, r->a1, an);
  for (int k = 0; k < 3; k++) { /* C:999-1011 */
    if (r->a1[k] > an[k]) { Imin[k] = an[k]; Imax[k] = r->a1[k]; }
    else { Imin[k] = r->a1[k]; Imax[k] = an[k]; }
  }
  r->amin = FN(max3)(Imin[0], Imin[1], Imin[2]);
  r->amax = FN(min3)(Imax[0], Imax[1], Imax[2]);
  r->hit = (r->amin < r->amax);
  if (!r->hit) return;
  for (int k = 0; k < 3; k++) { /* C:750-760, 762-767, 769-774 */
    if (u0[k] != 0) p[k] = mh->p[k] / u0[k];
    else p[k] = TGO_INF;
    r->pabs[k] = fabs(p[k]);
    r->update[k] = FN(sign)(u0[k]);
  }
  for (int k = 0; k < 3; k++) { /* C:776-788 */
    CT e = M[k] + r->amin * u0[k];
    CT ep = e - mh->min[k];
    r->iv[k] = (int)(unsigned int)(long long)(ep / mh->p[k]);
    r->iv[k] -= (unsigned int)(long long)(ep / mh->d[k]);
  }
  for (int k = 0; k < 3; k++) { /* C:790-802 */
    if (r->update[k] == 1) next[k] = r->iv[k] + 1;
    else if (r->update[k] == -1) next[k] = r->iv[k];
    else next[k] = TGO_INF * n[k];
  }
  for (int k = 0; k < 3; k++) /* C:804-809 */
    r->D[k] = ((CT)next[k]) * p[k] + r->a1[k] - r->amin;
  r->ac = r->amin;
}

/* C:948-953 */
static inline int FN(in_map)(const unsigned int n[3], const int iv[3])
{
  return (iv[0] >= 0 && (unsigned int)iv[0] < n[0] && iv[1] >= 0 && (unsigned int)iv[1] < n[1] &&
          iv[2] >= 0 && (unsigned int)iv[2] < n[2]);
}

/* C:956-961 with C:1034-1041 (squared distance, strict test against 1). */
static inline int FN(outside_sun)(const CT M[3], const CT u0[3], CT ac)
{
  CT d;
  d = (M[0] + ac * u0[0]) * (M[0] + ac * u0[0]);
  d += (M[1] + ac * u0[1]) * (M[1] + ac * u0[1]);
  d += (M[2] + ac * u0[2]) * (M[2] + ac * u0[2]);
  return d > 1;
}

/* C:811-912: one Siddon step.  Returns the segment length through the voxel
 * that was current on entry (*cur receives its index) and advances the state.
 * Returns 0 in *ok when no branch of the cascade applies (NaN distances: the
 * reference would spin forever; the oracle reports it instead). */
static inline CT FN(voxel_step)(FN(ray) *r, int cur[3], int *ok)
{
  CT *D = r->D;
  CT d;
  cur[0] = r->iv[0]; cur[1] = r->iv[1]; cur[2] = r->iv[2];
  *ok = 1;
  if ((D[0] < D[1]) && (D[0] < D[2])) {
    d = D[0]; r->ac += d;
    r->iv[0] += r->update[0];
    D[1] -= D[0]; D[2] -= D[0]; D[0] = r->pabs[0];
  } else if ((D[1] < D[0]) && (D[1] < D[2])) {
    d = D[1]; r->ac += d;
    r->iv[1] += r->update[1];
    D[0] -= D[1]; D[2] -= D[1]; D[1] = r->pabs[1];
  } else if ((D[2] < D[0]) && (D[2] < D[1])) {
    d = D[2]; r->ac += d;
    r->iv[2] += r->update[2];
    D[0] -= D[2]; D[1] -= D[2]; D[2] = r->pabs[2];
  } else if ((D[0] == D[1]) && (D[0] < D[2])) {
    d = D[0]; r->ac += d;
    r->iv[0] += r->update[0]; r->iv[1] += r->update[1];
    D[2] -= D[0]; D[0] = r->pabs[0]; D[1] = r->pabs[1];
  } else if ((D[1] == D[2]) && (D[1] < D[0])) {
    d = D[1]; r->ac += d;
    r->iv[1] += r->update[1]; r->iv[2] += r->update[2];
    D[0] -= D[1]; D[1] = r->pabs[1]; D[2] = r->pabs[2];
  } else if ((D[0] == D[2]) && (D[0] < D[1])) {
    d = D[0]; r->ac += d;
    r->iv[0] += r->update[0]; r->iv[2] += r->update[2];
    D[1] -= D[0]; D[0] = r->pabs[0]; D[2] = r->pabs[2];
  } else if ((D[0] == D[1]) && (D[0] == D[2])) {
    d = D[0]; r->ac += d;
    r->iv[0] += r->update[0]; r->iv[1] += r->update[1]; r->iv[2] += r->update[2];
    D[0] = r->pabs[0]; D[1] = r->pabs[1]; D[2] = r->pabs[2];
  } else {
    *ok = 0;
    d = 0;
  }
  return d;
}

/* Unit vector of pixel (i, j) of one image: C:516, 522-523. */
static inline void FN(pixel_ray)(const FN(ihdr) *ih, unsigned int i, unsigned int j, CT u0[3])
{
  CT gamma = FN(pixel2physical)(i, ih->s[0], ih->p[0], ih->v[0]);
  CT lambda = FN(pixel2physical)(j, ih->s[1], ih->p[1], ih->v[1]);
  FN(rotated_unit_vector)(lambda, gamma, ih->R, u0);
}

/* C:500-528 + C:530-580 + C:915-944 for one image t.  Returns the number of
 * Siddon steps taken (for step counting), or -1 on a NaN stall. */
static long long FN(image_projector)(const tgo_map *m, const tgo_data *dd, CT *cube, CT *data,
                                     int kind, int sun, int t, const FN(mhdr) *mh, const FN(ihdr) *ih)
{
  const unsigned int n[3] = {(unsigned)m->n[0], (unsigned)m->n[1], (unsigned)m->n[2]};
  long long steps = 0;
  const long long tbin = dd->tindex ? dd->tindex[t] : t;
  const long long toff = (kind == TGO_PJT || kind == TGO_BPJT) ? tbin * m->s[3] : 0;
  for (unsigned int i = 0; i < (unsigned)dd->n1; i++) {
    for (unsigned int j = 0; j < (unsigned)dd->n2; j++) {
      if (dd->mask && dd->mask[i * dd->ms[0] + j * dd->ms[1] + (long long)t * dd->ms[2]] != 0)
        continue;
      CT u0[3];
      FN(ray) r;
      FN(pixel_ray)(ih, i, j, u0);
      FN(ray_setup)(mh, n, ih->M, u0, &r);
      if (!r.hit) continue;
      const long long did = i * dd->s[0] + j * dd->s[1] + (long long)t * dd->s[2];
      while (FN(in_map)(n, r.iv) && (!sun || FN(outside_sun)(ih->M, u0, r.ac))) {
        int cur[3], ok;
        CT d = FN(voxel_step)(&r, cur, &ok);
        if (!ok) return -1;
        steps++;
        const long long vid = cur[0] * m->s[0] + cur[1] * m->s[1] + cur[2] * m->s[2] + toff;
        if (kind == TGO_PJ || kind == TGO_PJT) {
          if (data) {
#pragma omp atomic
            data[did] += d * cube[vid];
          }
        } else if (kind == TGO_BPJ || kind == TGO_BPJT) {
#pragma omp atomic
          cube[vid] += d * data[did];
        }
      }
    }
  }
  return steps;
}

/* C:460-498: the OpenMP team splits the images; t0/t1 select a view range
 * (t1 = t0 + 1 restates conic_image_projector, C:207-248). */
int FN(tgo_siddon)(const tgo_map *m, const tgo_data *dd, CT *cube, CT *data, int kind, int sun,
                   int nthread, int t0, int t1, long long *steps_out)
{
  FN(mhdr) mh;
  FN(ihdr) *ih;
  long long total = 0;
  int err = 0;
  if (t1 < 0) t1 = dd->nv;
  if (t0 < 0 || t1 > dd->nv || t0 > t1) return -2;
  ih = (FN(ihdr) *)malloc(sizeof(FN(ihdr)) * (size_t)(dd->nv > 0 ? dd->nv : 1));
  if (!ih) return -3;
  for (int t = 0; t < dd->nv; t++) FN(load_image)(dd->hdr + (size_t)t * TGO_HDR, &ih[t]);
  FN(load_map)(m, &mh);
  if (nthread <= 0) nthread = omp_get_max_threads();
#pragma omp parallel for num_threads(nthread) schedule(static) reduction(+ : total) reduction(| : err)
  for (int t = t0; t < t1; t++) {
    long long s = FN(image_projector)(m, dd, cube, data, kind, sun, t, &mh, &ih[t]);
    if (s < 0) err = 1;
    else total += s;
  }
  free(ih);
  if (steps_out) *steps_out = total;
  return err ? -1 : 0;
}

/* C:250-307: one ray (i, j, t).  Writes the visited voxel indices and lengths
 * (up to cap entries) and returns the step count, -1 on NaN stall. */
long long FN(tgo_ray_trace)(const tgo_map *m, const double *hdr18, int i, int j, int sun,
                            int *iv_out, CT *len_out, CT *u_out, long long cap)
{
  FN(mhdr) mh;
  FN(ihdr) ih;
  const unsigned int n[3] = {(unsigned)m->n[0], (unsigned)m->n[1], (unsigned)m->n[2]};
  CT u0[3];
  FN(ray) r;
  long long steps = 0;
  FN(load_map)(m, &mh);
  FN(load_image)(hdr18, &ih);
  FN(pixel_ray)(&ih, (unsigned)i, (unsigned)j, u0);
  if (u_out) { u_out[0] = u0[0]; u_out[1] = u0[1]; u_out[2] = u0[2]; }
  FN(ray_setup)(&mh, n, ih.M, u0, &r);
  if (!r.hit) return 0;
  while (FN(in_map)(n, r.iv) && (!sun || FN(outside_sun)(ih.M, u0, r.ac))) {
    int cur[3], ok;
    CT d = FN(voxel_step)(&r, cur, &ok);
    if (!ok) return -1;
    if (steps < cap) {
      if (iv_out) { iv_out[3 * steps] = cur[0]; iv_out[3 * steps + 1] = cur[1]; iv_out[3 * steps + 2] = cur[2]; }
      if (len_out) len_out[steps] = d;
    }
    steps++;
  }
  return steps;
}

/* C:582-648: unit vector per pixel; pixels whose data value is NaN are
 * skipped (C:638).  u has shape (n1, n2, nv, 3), C-contiguous. */
void FN(tgo_full_unit_vector)(const tgo_data *dd, const CT *data, CT *u)
{
  for (int t = 0; t < dd->nv; t++) {
    FN(ihdr) ih;
    FN(load_image)(dd->hdr + (size_t)t * TGO_HDR, &ih);
    for (unsigned int i = 0; i < (unsigned)dd->n1; i++)
      for (unsigned int j = 0; j < (unsigned)dd->n2; j++) {
        if (data) {
          CT x = data[i * dd->s[0] + j * dd->s[1] + (long long)t * dd->s[2]];
          if (x != x) continue;
        }
        CT u0[3];
        FN(pixel_ray)(&ih, i, j, u0);
        size_t o = (((size_t)i * dd->n2 + j) * dd->nv + t) * 3;
        u[o] = u0[0]; u[o + 1] = u0[1]; u[o + 2] = u0[2];
      }
  }
}

/* C:650-720: a1/an per pixel from a unit-vector array shaped like above. */
void FN(tgo_full_intersection_parameters)(const tgo_map *m, const tgo_data *dd, const CT *u, CT *a1, CT *an)
{
  FN(mhdr) mh;
  FN(load_map)(m, &mh);
  for (int t = 0; t < dd->nv; t++) {
    FN(ihdr) ih;
    FN(load_image)(dd->hdr + (size_t)t * TGO_HDR, &ih);
    for (unsigned int i = 0; i < (unsigned)dd->n1; i++)
      for (unsigned int j = 0; j < (unsigned)dd->n2; j++) {
        size_t o = (((size_t)i * dd->n2 + j) * dd->nv + t) * 3;
        FN(intersection_parameters)(&mh, ih.M, u + o, a1 + o, an + o);
      }
  }
}

/* Per-ray step count (value independent), used to pin the GPU count kernel. */
void FN(tgo_ray_steps)(const tgo_map *m, const tgo_data *dd, int sun, unsigned int *out)
{
  FN(mhdr) mh;
  const unsigned int n[3] = {(unsigned)m->n[0], (unsigned)m->n[1], (unsigned)m->n[2]};
  FN(load_map)(m, &mh);
#pragma omp parallel for schedule(dynamic, 1)
  for (int t = 0; t < dd->nv; t++) {
    FN(ihdr) ih;
    FN(load_image)(dd->hdr + (size_t)t * TGO_HDR, &ih);
    for (unsigned int i = 0; i < (unsigned)dd->n1; i++)
      for (unsigned int j = 0; j < (unsigned)dd->n2; j++) {
        size_t o = ((size_t)t * dd->n1 + i) * dd->n2 + j;
        out[o] = 0;
        if (dd->mask && dd->mask[i * dd->ms[0] + j * dd->ms[1] + (long long)t * dd->ms[2]] != 0)
          continue;
        CT u0[3];
        FN(ray) r;
        FN(pixel_ray)(&ih, i, j, u0);
        FN(ray_setup)(&mh, n, ih.M, u0, &r);
        if (!r.hit) continue;
        unsigned int c = 0;
        while (FN(in_map)(n, r.iv) && (!sun || FN(outside_sun)(ih.M, u0, r.ac))) {
          int cur[3], ok;
          FN(voxel_step)(&r, cur, &ok);
          if (!ok) { c = 0xffffffffu; break; }
          c++;
        }
        out[o] = c;
      }
  }
}

#undef FN
#undef CAT
#undef CAT_
