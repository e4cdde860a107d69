/*
 * CPU oracle (test infrastructure only) -- restatement of STIR's
 * ray-tracing projection matrix, its symmetries and the OSEM update.
 * See pmrt_oracle.h for scope and the list of pins.
 *
 * Build: gcc -O2 -std=c11 -ffp-contract=off -fopenmp -fPIC -shared
 * (no -ffast-math: the float evaluation order below mirrors the reference
 * and must not be re-associated or contracted).
 */
#include "pmrt_oracle.h"

#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define OR_PI 3.14159265358979323846264338327950288

/* ------------------------------------------------------------------ */
/* small helpers with the reference's exact semantics                  */
/* ------------------------------------------------------------------ */

/* stir::round(float) -- round.inl:59-66 */
static int round_f(float x) { return x >= 0 ? (int)(x + 0.5F) : -(int)(-x + 0.5F); }
/* std::min / std::max on floats (first argument wins ties) */
static float minf_(float a, float b) { return (b < a) ? b : a; }
static float maxf_(float a, float b) { return (a < b) ? b : a; }
static int mini_(int a, int b) { return (b < a) ? b : a; }
/* stir::modulo(float,float) -- modulo.h:52-79 */
static float modulo_f(float a, float b)
{
  const double r = fmod((double)a, (double)b);
  float res = (float)(r < 0 ? r + fabs((double)b) : r);
  const float abs_b = b >= 0 ? b : -b;
  if (res >= abs_b) res -= abs_b;
  return res;
}
/* norm() of a float 3-vector: float inner product then double sqrt
 * (BasicCoordinate.inl:408-429; order z,y,x = [1],[2],[3]) */
static double norm3(float z, float y, float x)
{
  float acc = 0.F;
  acc = acc + z * z;
  acc = acc + y * y;
  acc = acc + x * x;
  return sqrt((double)acc);
}
/* RayTraceVoxelsOnCartesianGrid.cxx:45-49 */
static int is_half_integer(float a) { return fabsf(floorf(a) + .5F - a) < .0001F; }

/* ------------------------------------------------------------------ */
/* row container (ProjMatrixElemsForOneBin)                           */
/* ------------------------------------------------------------------ */
typedef struct { int16_t z, y, x; float w; } elem;
/* tw: TOF weights of a basic row, n x num_tof_bins (element value times TOF
 * kernel value, 0 where the reference drops the element); NULL for non-TOF */
typedef struct { elem* e; int n, cap; float* tw; } row;

static void row_reserve(row* r, int cap)
{
  if (cap <= r->cap) return;
  int nc = r->cap ? r->cap : 16;
  while (nc < cap) nc *= 2;
  r->e = (elem*)realloc(r->e, (size_t)nc * sizeof(elem));
  r->cap = nc;
}
static void row_push(row* r, int z, int y, int x, float w)
{
  if (r->n == r->cap) row_reserve(r, r->n + 1);
  r->e[r->n].z = (int16_t)z; r->e[r->n].y = (int16_t)y; r->e[r->n].x = (int16_t)x; r->e[r->n].w = w;
  r->n++;
}
static void row_free(row* r) { free(r->e); free(r->tw); r->e = NULL; r->tw = NULL; r->n = r->cap = 0; }
static void row_copy(row* dst, const row* src)
{
  row_reserve(dst, src->n);
  memcpy(dst->e, src->e, (size_t)src->n * sizeof(elem));
  dst->n = src->n;
}
/* coordinates_less: z, then y, then x (ProjMatrixElemsForOneBinValue.inl) */
static int coord_cmp(const void* a, const void* b)
{
  const elem* p = (const elem*)a; const elem* q = (const elem*)b;
  if (p->z != q->z) return p->z < q->z ? -1 : 1;
  if (p->y != q->y) return p->y < q->y ? -1 : 1;
  if (p->x != q->x) return p->x < q->x ? -1 : 1;
  return 0;
}
static void row_sort(row* r) { if (r->n > 1) qsort(r->e, (size_t)r->n, sizeof(elem), coord_cmp); }
/* ProjMatrixElemsForOneBin::merge -- ProjMatrixElemsForOneBin.cxx:162-277:
 * sorts both operands, then a sorted union where equal coordinates sum
 * (this.value + other.value). Returns early (no sort) if other is empty. */
static void row_merge(row* lor, row* lor2)
{
  if (lor2->n == 0) return;
  row_sort(lor);
  row_sort(lor2);
  const int out_cap = lor->n + lor2->n;
  elem* out = (elem*)malloc((size_t)out_cap * sizeof(elem));
  int i = 0, j = 0, k = 0;
  while (j < lor2->n) {
    if (i == lor->n) { while (j < lor2->n) out[k++] = lor2->e[j++]; break; }
    const int c = coord_cmp(&lor2->e[j], &lor->e[i]);
    if (c == 0) { out[k] = lor->e[i]; out[k].w = lor->e[i].w + lor2->e[j].w; k++; i++; j++; }
    else if (c < 0) { out[k++] = lor2->e[j++]; }
    else { out[k++] = lor->e[i++]; }
  }
  while (i < lor->n) out[k++] = lor->e[i++];
  free(lor->e);
  lor->e = out; lor->n = k; lor->cap = out_cap;
}

/* ------------------------------------------------------------------ */
/* plan                                                               */
/* ------------------------------------------------------------------ */
typedef struct {
  int inplane; /* 0 id,1 x->-x,2 y->-y,3 both,4 swap,5 (x,y)->(-y,x),6 (y,-x),7 (-y,-x) */
  int zsign;   /* +1 or -1 */
  int zadd;    /* z' = zsign*z + zadd */
} symop;

struct or_plan {
  or_geometry g;
  int nseg;
  /* projection geometry */
  float dphi, phi_offset, dalpha, radius, ring_spacing;
  int max_tang;
  float ax_samp[OR_MAX_SEGMENTS], m_offset[OR_MAX_SEGMENTS], delta[OR_MAX_SEGMENTS];
  int64_t seg_offset[OR_MAX_SEGMENTS]; /* STIR storage offset per segment index */
  int64_t num_bins;
  /* image */
  int max_z, max_y, max_x;
  /* symmetries */
  int sym90, sym180, symseg, syms, symz;
  int actual_bounds; /* use_actual_detector_boundaries after the set-up's checks */
  int ppa[OR_MAX_SEGMENTS];
  int ppr;
  float ap2z[OR_MAX_SEGMENTS];
  /* basic row cache: index (seg-min)*nviews*(2T+1) + view*(2T+1) + (tang+T), where
   * T = max(-min_tang, max_tang): the basic bin of a tangential position
   * may lie outside the data's range (e.g. -n/2 -> +n/2) and is still traced */
  int tmax;
  row* basic;
  /* time of flight */
  int ntof;
  float z_offset;  /* scanner z of image plane 0 subtracted in apply_tof_kernel */
  double* erf_tab;  /* FastErf samples */
  int tofsens;      /* "use time-of-flight sensitivities" (default 0) */
  int nt_sym[4];    /* 90, 180, swap segment, swap s before the TOF restriction */
  /* objective function and OSMAPOSL settings */
  const float* mult;        /* multiplicative factors per bin (caller-owned), NULL = 1 */
  int no_subset_sens;       /* "use subset sensitivities" := 0 */
  int map_multiplicative;   /* "MAP_model" := multiplicative (default additive) */
  float min_rel, max_rel;   /* "minimum/maximum relative change" */
  int prior_type;           /* 0 none, 1 quadratic, 2 relative difference */
  float beta, gamma, eps;   /* penalisation factor; RDP gamma and epsilon */
  int only_2D;
  const float* kappa;       /* caller-owned image or NULL */
  float w27[27];
};

#define SEGI(p, s) ((s) - (p)->g.min_segment_num)

static float get_axial_sampling(const or_plan* p, int seg)
{
  const int i = SEGI(p, seg);
  /* ProjDataInfoCylindrical.inl:121-129 */
  return p->ring_spacing / (p->g.max_ring_diff[i] != p->g.min_ring_diff[i] ? 2 : 1);
}

/* get_s -- ProjDataInfoCylindricalNoArcCorr.inl:79-82 (double libm sin, see header),
 * ProjDataInfoCylindricalArcCorr.inl:31-34 */
static float get_s(const or_plan* p, int tang)
{
  if (p->g.arc_corrected) return tang * p->g.bin_size;
  return (float)((double)p->radius * sin((double)(tang * p->dalpha)));
}
/* ProjDataInfoCylindrical.inl:65-68 */
static float get_phi(const or_plan* p, int view) { return view * p->dphi + p->phi_offset; }
/* ProjDataInfoCylindrical.inl:85-93 */
static float get_tantheta(const or_plan* p, int seg, int tang)
{
  const float delta = p->delta[SEGI(p, seg)];
  if (fabsf(delta) < 0.0001F) return 0;
  const float R = p->radius;
  const float s = get_s(p, tang);
  return (float)((delta * p->ring_spacing) / (2 * sqrt((double)(R * R - s * s))));
}
/* ProjDataInfo.inl:200-203 (double sqrt) */
static float get_costheta(const or_plan* p, int seg, int tang)
{
  const float tt = get_tantheta(p, seg, tang);
  return (float)(1 / sqrt((double)(1 + tt * tt)));
}
/* ProjDataInfoCylindrical.inl:71-82 */
static float get_m(const or_plan* p, int seg, int ax)
{
  return ax * get_axial_sampling(p, seg) - p->m_offset[SEGI(p, seg)];
}
static float get_t(const or_plan* p, int seg, int ax, int tang)
{
  return get_m(p, seg, ax) * get_costheta(p, seg, tang);
}
/* ProjDataInfo.cxx:110-115 and ProjDataInfoCylindricalArcCorr.h:70 */
static float get_sampling_in_s(const or_plan* p, int tang)
{
  if (p->g.arc_corrected) return p->g.bin_size;
  return fabsf(get_s(p, tang + 1) - get_s(p, tang - 1)) / 2;
}

/* ------------------------------------------------------------------ */
/* Siddon tracer -- RayTraceVoxelsOnCartesianGrid.cxx:52-269           */
/* ------------------------------------------------------------------ */
static void ray_trace(row* lor, float sz, float sy, float sx, float ez, float ey, float ex,
                      float vz, float vy, float vx, float nc)
{
  const float dz = ez - sz, dy = ey - sy, dx = ex - sx;
  if (norm3(dz, dy, dx) <= .00001F) return;
  const float d12 = (float)(norm3(dz * vz, dy * vy, dx * vx) * (double)nc);
  const int sign_x = dx >= 0 ? 1 : -1;
  const int sign_y = dy >= 0 ? 1 : -1;
  const int sign_z = dz >= 0 ? 1 : -1;
  const float small_difference = 1.E-4F;
  const int zdx = fabsf(dx) <= small_difference;
  const int zdy = fabsf(dy) <= small_difference;
  const int zdz = fabsf(dz) <= small_difference;
  {
    float iz = 0, iy = 0, ix = 0;
    if (zdz && is_half_integer(sz)) iz = .5F;
    else if (zdy && is_half_integer(sy)) iy = .5F;
    else if (zdx && is_half_integer(sx)) ix = .5F;
    if (norm3(iz, iy, ix) > .1) {
      ray_trace(lor, sz - iz, sy - iy, sx - ix, ez - iz, ey - iy, ex - ix, vz, vy, vx, nc / 2);
      ray_trace(lor, sz + iz, sy + iy, sx + ix, ez + iz, ey + iy, ex + ix, vz, vy, vx, nc / 2);
      row_sort(lor);
      return;
    }
  }
  const float inc_x = zdx ? d12 * 1000000.F : d12 / fabsf(dx);
  const float inc_y = zdy ? d12 * 1000000.F : d12 / fabsf(dy);
  const float inc_z = zdz ? d12 * 1000000.F : d12 / fabsf(dz);
  const float xmax = round_f(ex) + sign_x * 0.5F;
  const float ymax = round_f(ey) + sign_y * 0.5F;
  const float zmax = round_f(ez) + sign_z * 0.5F;
  const float axend = zdx ? d12 * 1000000.F : (xmax - sx) * inc_x * sign_x * .9999F;
  const float ayend = zdy ? d12 * 1000000.F : (ymax - sy) * inc_y * sign_y * .9999F;
  const float azend = zdz ? d12 * 1000000.F : (zmax - sz) * inc_z * sign_z * .9999F;
  const float amax = minf_(axend, minf_(ayend, azend));
  int cz = round_f(sz), cy = round_f(sy), cx = round_f(sx);
  float az = zdz ? -inc_z : ((cz - sz) - sign_z * 0.5F) * inc_z * sign_z;
  float ax = zdx ? -inc_x : ((cx - sx) - sign_x * 0.5F) * inc_x * sign_x;
  float ay = zdy ? -inc_y : ((cy - sy) - sign_y * 0.5F) * inc_y * sign_y;
  float a = maxf_(ax, maxf_(ay, az));
  if (zdx) ax = axend; else ax += inc_x;
  if (zdy) ay = ayend; else ay += inc_y;
  if (zdz) az = azend; else az += inc_z;
  while (a < amax) {
    if (ax < ay) {
      if (ax < az) { row_push(lor, cz, cy, cx, ax - a); a = ax; ax += inc_x; cx += sign_x; }
      else { row_push(lor, cz, cy, cx, az - a); a = az; az += inc_z; cz += sign_z; }
    } else if (ay < az) { row_push(lor, cz, cy, cx, ay - a); a = ay; ay += inc_y; cy += sign_y; }
    else { row_push(lor, cz, cy, cx, az - a); a = az; az += inc_z; cz += sign_z; }
  }
}

/* ray_trace_one_lor -- ProjMatrixByBinUsingRayTracing.cxx:438-579 */
static void ray_trace_one_lor(row* lor, float s, float t, float cphi, float sphi, float costheta,
                              float tantheta, float offset_in_z, float fovrad, float vz, float vy,
                              float vx, int restrict_cyl, int num_lors)
{
  float max_a, min_a;
  if (restrict_cyl) {
    if (fabsf(s) > fovrad) return;
    if (fabsf(s) == fovrad) { max_a = min_a = 0; }
    else { max_a = sqrtf(fovrad * fovrad - s * s); min_a = -max_a; }
  } else {
    if (fabsf(cphi) < 1.E-3 || fabsf(sphi) < 1.E-3) {
      if (fovrad < fabsf(s)) return;
      max_a = fovrad; min_a = -fovrad;
    } else {
      const int sgs = sphi < 0 ? -1 : 1, sgc = cphi < 0 ? -1 : 1;
      max_a = minf_((fovrad * sgs - s * cphi) / sphi, (fovrad * sgc + s * sphi) / cphi);
      min_a = maxf_((-fovrad * sgs - s * cphi) / sphi, (-fovrad * sgc + s * sphi) / cphi);
      if (min_a > max_a - 1.E-3 * vx) return;
    }
  }
  const float st_x = (s * cphi + max_a * sphi) / vx;
  const float st_y = (s * sphi - max_a * cphi) / vy;
  const float st_z = (t / costheta + offset_in_z - max_a * tantheta) / vz;
  const float sp_x = (s * cphi + min_a * sphi) / vx;
  const float sp_y = (s * sphi - min_a * cphi) / vy;
  const float sp_z = (t / costheta + offset_in_z - min_a * tantheta) / vz;
  const int fwd = st_z < sp_z || (st_z == sp_z && (st_y < sp_y || (st_y == sp_y && st_x <= sp_x)));
  const float nc = 1 / vx / num_lors;
  if (fwd) ray_trace(lor, st_z, st_y, st_x, sp_z, sp_y, sp_x, vz, vy, vx, nc);
  else ray_trace(lor, sp_z, sp_y, sp_x, st_z, st_y, st_x, vz, vy, vx, nc);
}

/* add_adjacent_z -- ProjMatrixByBinUsingRayTracing.cxx:832-954 */
static void add_adjacent_z(row* lor, float zf, float right)
{
  const int org = lor->n;
  const int z0 = lor->e[0].z;
  for (int zi = 1;; ++zi) {
    const float ov = minf_(right, zf + zi + .5F) - maxf_(0.F, zf + zi - .5F);
    if (ov <= 0.0001) break;
    const int nz = z0 + zi;
    row_reserve(lor, lor->n + org);
    for (int k = 0; k < org; ++k) {
      const elem e = lor->e[k];
      row_push(lor, nz, e.y, e.x, ov > .9999 ? e.w : e.w * ov);
    }
  }
  if (zf > 0.5) {
    const float ov = zf - .5F;
    const int nz = z0 - 1;
    row_reserve(lor, lor->n + org);
    for (int k = 0; k < org; ++k) {
      const elem e = lor->e[k];
      row_push(lor, nz, e.y, e.x, ov > .9999 ? e.w : e.w * ov);
    }
  }
  {
    const float ov = minf_(right, zf + .5F) - maxf_(0.F, zf - .5F);
    if (ov < .9999)
      for (int k = 0; k < org; ++k) lor->e[k].w *= ov;
  }
}

/* phi and s of the LOR between the two detectors of (view, tang) --
 * ProjMatrixByBinUsingRayTracing.cxx:613-634, with the detector pair of
 * ProjDataInfoCylindricalNoArcCorr::initialise_uncompressed_view_tangpos_to_det1det2
 * (ProjDataInfoCylindricalNoArcCorr.cxx:164-200; >> of a negative int rounds
 * towards minus infinity); both evaluated in double and rounded to float. */
static void actual_boundary_lor(const or_plan* p, int view, int tang, float* s, float* phi)
{
  const int ndet = p->g.num_detectors_per_ring;
  const int det1 = (view + (tang >> 1) + ndet) % ndet;
  const int det2 = (view - ((tang + 1) >> 1) + ndet / 2) % ndet;
  *phi = (float)((det1 + det2) * OR_PI / ndet - OR_PI / 2 + p->phi_offset);
  *s = (float)(p->g.effective_ring_radius * sin((det1 - det2) * OR_PI / ndet + OR_PI / 2));
}

/* calculate_proj_matrix_elems_for_one_bin -- ProjMatrixByBinUsingRayTracing.cxx:582-829 */
static int calc_row(const or_plan* p, int seg, int view, int ax, int tang, row* lor)
{
  const or_geometry* g = &p->g;
  lor->n = 0;
  float s_in_mm = get_s(p, tang);
  float phi = get_phi(p, view);
  if (p->actual_bounds) actual_boundary_lor(p, view, tang, &s_in_mm, &phi);
  const float cphi = cosf(phi);
  const float sphi = sinf(phi);
  const float tantheta = get_tantheta(p, seg, tang);
  const float costheta = 1 / sqrtf(1 + tantheta * tantheta);
  const float t_in_mm = get_t(p, seg, ax, tang);
  const float samp_z = (get_axial_sampling(p, seg) * get_costheta(p, seg, tang)) / costheta;
  const int num_lors = (int)ceil(samp_z / g->vox_z - 1.E-3);
  if (fabs(samp_z / g->vox_z - num_lors) > 1E-3) return -1;
  const float zfirst = (-samp_z / (2 * num_lors) * (num_lors - 1)) - g->origin_z;
  float offset_in_z = zfirst + (p->max_z + g->min_z) / 2.F * g->vox_z;
  if (tantheta == 0) {
    if (fabs(modulo_f((t_in_mm + offset_in_z) / g->vox_z, 1.F) - .5) < .001)
      offset_in_z -= .1F * g->vox_z;
  }
  const float fovrad = minf_((mini_(p->max_x, -g->min_x)) * g->vox_x, (mini_(p->max_y, -g->min_y)) * g->vox_y);
  if (g->num_tangential_LORs == 1) {
    ray_trace_one_lor(lor, s_in_mm, t_in_mm, cphi, sphi, costheta, tantheta, offset_in_z, fovrad,
                      g->vox_z, g->vox_y, g->vox_x, g->restrict_to_cylindrical_FOV, num_lors);
  } else {
    row tmp = {0};
    const int n = g->num_tangential_LORs;
    /* twice the spacing with actual detector boundaries (:760-762) */
    const float s_inc = (p->actual_bounds ? 2 : 1) * get_sampling_in_s(p, tang) / n;
    float cur = s_in_mm - s_inc * (n - 1) / 2.F;
    for (int k = 1; k <= n; ++k, cur += s_inc) {
      tmp.n = 0;
      ray_trace_one_lor(&tmp, cur, t_in_mm, cphi, sphi, costheta, tantheta, offset_in_z, fovrad,
                        g->vox_z, g->vox_y, g->vox_x, g->restrict_to_cylindrical_FOV, num_lors * n);
      row_merge(lor, &tmp);
    }
    row_free(&tmp);
  }
  if (lor->n > 0) {
    if (tantheta == 0) {
      const float zf = lor->e[0].z + g->origin_z / g->vox_z - (p->max_z + g->min_z) / 2.F;
      const float left = (t_in_mm - samp_z / 2) / g->vox_z;
      const float right = (t_in_mm + samp_z / 2) / g->vox_z;
      add_adjacent_z(lor, zf - left, right - left);
    } else if (num_lors > 1) {
      row nxt = {0};
      row_copy(&nxt, lor);
      for (int zi = 1; zi < num_lors; ++zi) {
        for (int k = 0; k < nxt.n; ++k) nxt.e[k].z = (int16_t)(nxt.e[k].z + 1);
        row_merge(lor, &nxt);
      }
      row_free(&nxt);
    }
  }
  return 0;
}

/* ------------------------------------------------------------------ */
/* symmetries -- DataSymmetriesForBins_PET_CartesianGrid.{cxx,inl},    */
/* SymmetryOperations_PET_CartesianGrid.inl                            */
/* ------------------------------------------------------------------ */

/* find_transform_z -- .inl:67-118 */
static int find_transform_z(const or_plan* p, int seg, int ax)
{
  const int i = SEGI(p, seg);
  const float delta = p->delta[i];
  const float tz = (2 * p->ppa[i] * (ax) + p->ppr * delta + 2 * p->ap2z[i]);
  return (int)floor(tz + 0.5);
}

enum { OP_ID = 0, OP_XMX = 1, OP_YMY = 2, OP_XMX_YMY = 3, OP_XY_YX = 4, OP_XMY_YX = 5, OP_XY_YMX = 6, OP_XMY_YMX = 7 };

static symop mkop(int inplane, int zq, int q, int z_shift)
{
  symop o;
  o.inplane = inplane;
  o.zsign = zq ? -1 : 1;
  o.zadd = zq ? q + z_shift : z_shift;
  return o;
}

/* find_sym_op_bin0 -- .inl:120-220 (cylindrical) */
static symop find_sym_op_bin0(const or_plan* p, int seg, int view, int ax)
{
  const int q = find_transform_z(p, abs(seg), p->symz ? 0 : ax);
  const int z_shift = p->symz ? p->ppa[SEGI(p, seg)] * ax : 0;
  const int v180 = p->g.num_views, v135 = v180 / 4 * 3, v90 = v180 / 2, v45 = v180 / 4;
  if (p->sym90 && view > v90 && view <= v135)
    return (!p->symseg || seg >= 0) ? mkop(OP_XMY_YX, 0, q, z_shift) : mkop(OP_XMY_YX, 1, q, z_shift);
  else if (p->sym90 && view > v45 && view <= v90)
    return (!p->symseg || seg >= 0) ? mkop(OP_XY_YX, 1, q, z_shift) : mkop(OP_XY_YX, 0, q, z_shift);
  else if (p->sym180 && view > v90)
    return (!p->symseg || seg >= 0) ? mkop(OP_XMX, 1, q, z_shift) : mkop(OP_XMX, 0, q, z_shift);
  else {
    if (p->symseg && seg < 0) return mkop(OP_ID, 1, q, z_shift);
    return mkop(OP_ID, 0, q, z_shift);
  }
}

/* find_sym_op_general_bin -- .inl:223-396 (cylindrical) */
static symop find_sym_op_general_bin(const or_plan* p, int s, int seg, int view, int ax)
{
  const int q = find_transform_z(p, abs(seg), p->symz ? 0 : ax);
  const int z_shift = p->symz ? p->ppa[SEGI(p, seg)] * ax : 0;
  const int v180 = p->g.num_views, v135 = v180 / 4 * 3, v90 = v180 / 2, v45 = v180 / 4;
  const int spos = !p->syms || s > 0;
  if (p->sym90 && view > v90 && view <= v135) {
    if (!p->symseg || seg > 0)
      return spos ? mkop(OP_XMY_YX, 0, q, z_shift) : mkop(OP_XY_YMX, 1, q, z_shift);
    else if (seg < 0)
      return spos ? mkop(OP_XMY_YX, 1, q, z_shift) : mkop(OP_XY_YMX, 0, q, z_shift);
    else
      return spos ? mkop(OP_XMY_YX, 0, q, z_shift) : mkop(OP_XY_YMX, 0, q, z_shift);
  } else if (p->sym90 && view > v45 && view <= v90) {
    if (!p->symseg || seg > 0)
      return spos ? mkop(OP_XY_YX, 1, q, z_shift) : mkop(OP_XMY_YMX, 0, q, z_shift);
    else if (seg < 0)
      return spos ? mkop(OP_XY_YX, 0, q, z_shift) : mkop(OP_XMY_YMX, 1, q, z_shift);
    else
      return spos ? mkop(OP_XY_YX, 0, q, z_shift) : mkop(OP_XMY_YMX, 0, q, z_shift);
  } else if (p->sym180 && view > v90) {
    if (!p->symseg || seg > 0)
      return spos ? mkop(OP_XMX, 1, q, z_shift) : mkop(OP_YMY, 0, q, z_shift);
    else
      return spos ? mkop(OP_XMX, 0, q, z_shift) : mkop(OP_YMY, 1, q, z_shift);
  } else {
    const int sneg = p->syms && s < 0;
    if (!p->symseg || seg > 0)
      return sneg ? mkop(OP_XMX_YMY, 1, q, z_shift) : mkop(OP_ID, 0, q, z_shift);
    else if (seg < 0)
      return sneg ? mkop(OP_XMX_YMY, 0, q, z_shift) : mkop(OP_ID, 1, q, z_shift);
    else
      return sneg ? mkop(OP_XMX_YMY, 0, q, z_shift) : mkop(OP_ID, 0, q, z_shift);
  }
}

/* find_basic_view_segment_numbers -- .inl:398-447 */
static int find_basic_vs(const or_plan* p, int* view, int* seg)
{
  int change = 0;
  const int nv = p->g.num_views, v90 = nv >> 1, v45 = v90 >> 1, v135 = v90 + v45;
  if (p->symseg && *seg < 0) { *seg = -*seg; change = 1; }
  if (p->sym90) {
    if (*view >= v135) { *view = nv - *view; return 1; }
    else if (*view >= v90) { *view = *view - v90; return 1; }
    else if (*view > v45) { *view = v90 - *view; return 1; }
  } else if (p->sym180) {
    if (*view > v90) { *view = nv - *view; return 1; }
  }
  return change;
}
/* find_basic_bin -- .inl:449-478 */
static void find_basic_bin(const or_plan* p, int* seg, int* view, int* ax, int* tang)
{
  find_basic_vs(p, view, seg);
  if (p->syms && *tang < 0) *tang = -*tang;
  if (p->symz && *ax != 0) *ax = 0;
}

static void apply_op(const symop* o, int z, int y, int x, int* oz, int* oy, int* ox)
{
  int nx = x, ny = y;
  switch (o->inplane) {
    case OP_ID: break;
    case OP_XMX: nx = -x; break;
    case OP_YMY: ny = -y; break;
    case OP_XMX_YMY: nx = -x; ny = -y; break;
    case OP_XY_YX: nx = y; ny = x; break;
    case OP_XMY_YX: nx = -y; ny = x; break;
    case OP_XY_YMX: nx = y; ny = -x; break;
    case OP_XMY_YMX: nx = -y; ny = -x; break;
  }
  *ox = nx; *oy = ny; *oz = o->zsign * z + o->zadd;
}

/* find_symmetry_operation_from_basic_bin -- .inl:664-674 */
static symop sym_op_for_bin(const or_plan* p, int* seg, int* view, int* ax, int* tang)
{
  symop o = (*tang == 0) ? find_sym_op_bin0(p, *seg, *view, *ax)
                         : find_sym_op_general_bin(p, *tang, *seg, *view, *ax);
  find_basic_bin(p, seg, view, ax, tang);
  return o;
}

/* get_related_view_segment_numbers -- .inl:800-849 */
static int related_vs(const or_plan* p, int view, int seg, int* rv, int* rs)
{
  const int nv = p->g.num_views;
  const int symz = p->symseg && seg != 0;
  int n = 0;
  rv[n] = view; rs[n++] = seg;
  if (symz) { rv[n] = view; rs[n++] = -seg; }
  if (p->sym180 && p->sym90 && (view % (nv / 2)) != nv / 4) {
    const int r = view < nv / 2 ? view + nv / 2 : view - nv / 2;
    rv[n] = r; rs[n++] = seg;
    if (symz) { rv[n] = r; rs[n++] = -seg; }
  }
  if (p->sym180 && (view % (nv / 2)) != 0) {
    rv[n] = nv - view; rs[n++] = seg;
    if (symz) { rv[n] = nv - view; rs[n++] = -seg; }
  }
  if (p->sym90 && (view % (nv / 4)) != 0) {
    const int r = (nv / 2 - view + nv) % nv;
    rv[n] = r; rs[n++] = seg;
    if (symz) { rv[n] = r; rs[n++] = -seg; }
  }
  return n;
}

/* ------------------------------------------------------------------ */
/* plan creation                                                       */
/* ------------------------------------------------------------------ */
/* ------------------------------------------------------------------ */
/* time of flight                                                       */
/* ------------------------------------------------------------------ */
/* FastErf (FastErf.inl:30-80): erf sampled at 200000+2 points on [-5, 5],
 * linear interpolation; ProjMatrixByBin.cxx:174-176 sets 200000 samples and
 * the maximum sample value keeps its default 5 (FastErf.h:57). */
#define OR_ERF_SAMPLES 200000
#define OR_ERF_MAX 5.0
static double erf_period(void) { return (2 * OR_ERF_MAX) / OR_ERF_SAMPLES; }
static double* erf_table(void)
{
  double* t = (double*)malloc(sizeof(double) * (OR_ERF_SAMPLES + 2));
  const double per = erf_period();
  for (int i = 0; i < OR_ERF_SAMPLES + 2; ++i) t[i] = erf(i * per - OR_ERF_MAX);
  return t;
}
static double erf_interp(const double* tab, double xp)
{
  xp = fmax(fmin(OR_ERF_MAX, xp), -OR_ERF_MAX);
  const double xi = (xp + OR_ERF_MAX) / erf_period();
  const int lo = (int)floor(xi);
  return tab[lo] + (xi - lo) * (tab[lo + 1] - tab[lo]);
}
/* ProjMatrixByBin::get_tof_value (ProjMatrixByBin.inl:186-195) */
static float tof_value(const or_plan* p, float d1, float d2)
{
  const float r = p->g.tof_r_sqrt2_gauss_sigma;
  const float d1_n = d1 * r, d2_n = d2 * r;
  if ((d1_n >= 4.F && d2_n >= 4.F) || (d1_n <= -4.F && d2_n <= -4.F)) return 0.F;
  return (float)(0.5 * (erf_interp(p->erf_tab, d2_n) - erf_interp(p->erf_tab, d1_n)));
}
/* to_0_2pi for floats: modulo(float, float) (modulo.h:62-76,136-141) */
static float to_0_2pi_f(float phi)
{
  const float b = (float)(2 * OR_PI);
  double r = fmod((double)phi, (double)b);
  r = r < 0 ? r + fabs((double)b) : r;
  float res = (float)r;
  if (res >= b) res -= b;
  return res;
}
/* Mid-point and unit direction (z, y, x) of a bin's LOR as apply_tof_kernel
 * forms them (ProjMatrixByBin.inl:117-137): get_LOR
 * (ProjDataInfoCylindrical.cxx:515-547), LORInAxialAndNoArcCorrSinogramCoordinates
 * (LORCoordinates.inl:106-121), LORInCylinderCoordinates and LORAs2Points
 * (:128-139,287-298). Float arithmetic; asin/cos/sin are the float overloads
 * (parity at the last bit unpinned: the overload the reference picks depends
 * on its header inclusion order). */
static void tof_lor(const or_plan* p, int seg, int view, int ax, int tang, float mid[3], float unit[3])
{
  const float s = get_s(p, tang), m = get_m(p, seg, ax), tt = get_tantheta(p, seg, tang), phi = get_phi(p, view);
  const float R = p->radius;
  const float max_a = sqrtf(R * R - s * s), min_a = -max_a;
  float z1 = m - max_a * tt, z2 = m - min_a * tt;
  float ph = to_0_2pi_f(phi), beta = asinf(s / R);
  int swapped = 0;
  if ((double)ph >= OR_PI) {
    ph -= (float)OR_PI;
    beta = -beta;
    const float tz = z1; z1 = z2; z2 = tz;
    swapped = 1;
  }
  const float psi1 = to_0_2pi_f(ph + beta), psi2 = to_0_2pi_f(ph - beta + (float)OR_PI);
  float p1[3] = {z1, -R * cosf(psi1), R * sinf(psi1)};
  float p2[3] = {z2, -R * cosf(psi2), R * sinf(psi2)};
  if (swapped)
    for (int i = 0; i < 3; ++i) { const float t = p1[i]; p1[i] = p2[i]; p2[i] = t; }
  float diff[3];
  for (int i = 0; i < 3; ++i) {
    mid[i] = (p1[i] + p2[i]) * 0.5F;
    diff[i] = p2[i] - mid[i];
  }
  const float n = (float)norm3(diff[0], diff[1], diff[2]);
  for (int i = 0; i < 3; ++i) unit[i] = diff[i] / n;
}
/* apply_tof_kernel (ProjMatrixByBin.inl:115-181) on a basic row: TOF weights
 * value * kernel for every TOF bin, 0 where the kernel value is not positive. */
static void tof_weights(const or_plan* p, int seg, int view, int tang, row* r)
{
  const int nt = p->ntof;
  r->tw = (float*)malloc(sizeof(float) * (size_t)(r->n > 0 ? r->n : 1) * nt);
  float mid[3], unit[3];
  tof_lor(p, seg, view, 0, tang, mid, unit);
  for (int e = 0; e < r->n; ++e) {
    /* get_physical_coordinates_for_indices (DiscretisedDensity.inl:175-179,
     * DiscretisedDensityOnCartesianGrid.inl:104-112), minus the scanner offset */
    const float vz = (p->g.vox_z * (float)r->e[e].z + p->g.origin_z) - p->z_offset;
    const float vy = (p->g.vox_y * (float)r->e[e].y + p->g.origin_y) - 0.F;
    const float vx = (p->g.vox_x * (float)r->e[e].x + p->g.origin_x) - 0.F;
    float ip = 0.F;
    ip += (vz - mid[0]) * unit[0];
    ip += (vy - mid[1]) * unit[1];
    ip += (vx - mid[2]) * unit[2];
    const float d2 = -ip;
    for (int k = 0; k < nt; ++k) {
      const float v = tof_value(p, p->g.tof_low[k] - d2, p->g.tof_high[k] - d2);
      float w = 0.F;
      if (v > 0 && r->e[e].w != 0) w = r->e[e].w * v;
      r->tw[(size_t)e * nt + k] = w;
    }
  }
}

static size_t basic_index(const or_plan* p, int seg, int view, int tang)
{
  return ((size_t)SEGI(p, seg) * p->g.num_views + view) * (size_t)(2 * p->tmax + 1) + (size_t)(tang + p->tmax);
}

or_plan* or_plan_create(const or_geometry* g, char* err, int errlen)
{
  or_plan* p = (or_plan*)calloc(1, sizeof(or_plan));
  p->g = *g;
  p->min_rel = 0.F;      /* OSMAPOSLReconstruction.cxx:95-104 */
  p->max_rel = FLT_MAX;
  p->nseg = g->max_segment_num - g->min_segment_num + 1;
  if (p->nseg <= 0 || p->nseg > OR_MAX_SEGMENTS) { snprintf(err, errlen, "bad segment range"); free(p); return NULL; }
  /* ProjDataInfoCylindrical ctor -- ProjDataInfoCylindrical.cxx:56-117 */
  p->dphi = (float)(OR_PI / g->num_views);
  p->phi_offset = g->intrinsic_tilt;
  {
    const int ndet = g->num_detectors_per_ring;
    if (ndet > 2 && g->num_views * 2 != ndet && (ndet % (g->num_views * 2)) == 0) {
      const int mash = ndet / 2 / g->num_views;
      const float offset_inc = (float)(OR_PI / (ndet / 2) * (mash - 1) / 2.F);
      p->phi_offset += offset_inc;
    }
  }
  p->radius = g->effective_ring_radius;
  p->ring_spacing = g->ring_spacing;
  p->dalpha = (float)(OR_PI / g->num_detectors_per_ring); /* NoArcCorr.cxx:77-93 */
  p->max_tang = g->min_tangential_pos_num + g->num_tangential_poss - 1;
  p->tmax = p->max_tang > -g->min_tangential_pos_num ? p->max_tang : -g->min_tangential_pos_num;
  if (g->do_symmetry_swap_segment && g->min_segment_num != -g->max_segment_num) {
    snprintf(err, errlen, "segment symmetry needs a symmetric segment range");
    free(p);
    return NULL;
  }
  for (int s = g->min_segment_num; s <= g->max_segment_num; ++s) {
    const int i = SEGI(p, s);
    p->ax_samp[i] = get_axial_sampling(p, s);
    /* m_offset -- ProjDataInfoCylindrical.cxx:139-145 */
    p->m_offset[i] = ((g->num_axial_poss[i] - 1 + 0) * p->ax_samp[i]) / 2;
    /* get_average_ring_difference -- .inl:132-138 */
    p->delta[i] = (g->min_ring_diff[i] + g->max_ring_diff[i]) / 2.F;
  }
  /* storage offsets in segment order 0,+1,-1,... (ProjData.cxx:641-661) */
  {
    int64_t off = 0;
    int seq[OR_MAX_SEGMENTS], ns = 0;
    seq[ns++] = 0;
    for (int s = 1; ns < p->nseg; ++s) {
      if (s <= g->max_segment_num) seq[ns++] = s;
      if (-s >= g->min_segment_num) seq[ns++] = -s;
    }
    for (int k = 0; k < ns; ++k) {
      const int i = SEGI(p, seq[k]);
      p->seg_offset[i] = off;
      off += (int64_t)g->num_axial_poss[i] * g->num_views * g->num_tangential_poss;
    }
    p->num_bins = off;
  }
  p->max_z = g->min_z + g->nz - 1;
  p->max_y = g->min_y + g->ny - 1;
  p->max_x = g->min_x + g->nx - 1;
  /* symmetries -- DataSymmetriesForBins_PET_CartesianGrid.cxx:241-378 */
  p->sym90 = g->do_symmetry_90degrees_min_phi;
  p->sym180 = g->do_symmetry_90degrees_min_phi || g->do_symmetry_180degrees_min_phi;
  p->symseg = g->do_symmetry_swap_segment;
  p->syms = g->do_symmetry_swap_s;
  p->symz = g->do_symmetry_shift_z;
  /* use_actual_detector_boundaries is reset when the data are arc-corrected,
   * view-mashed or axially compressed (ProjMatrixByBinUsingRayTracing.cxx:300-327) */
  p->actual_bounds = g->use_actual_detector_boundaries && !g->arc_corrected &&
                     g->num_views * 2 == g->num_detectors_per_ring;
  for (int si = 0; si < p->nseg; ++si)
    if (g->min_ring_diff[si] != g->max_ring_diff[si]) p->actual_bounds = 0;
  p->ntof = g->num_tof_bins > 1 ? g->num_tof_bins : 1;
  if (p->ntof > OR_MAX_TOF) { snprintf(err, errlen, "too many TOF bins"); free(p); return NULL; }
  if (p->ntof > 1) {
    /* ProjMatrixByBin.inl:128-131 */
    p->z_offset = (g->num_rings - 1) / 2.0F * g->ring_spacing;
    p->erf_tab = erf_table();
  }
  {
    const float zo = g->origin_z / g->vox_z;
    if (fabsf(round_f(zo) - zo) > 1.E-3F) { snprintf(err, errlen, "z origin not a multiple of plane spacing"); free(p); return NULL; }
    if (fabsf(g->vox_y - g->vox_x) > 2.E-3F) p->sym90 = 0;
    if (g->num_views % 4 != 0) p->sym90 = 0;
    if (g->num_views % 2 != 0) p->sym180 = 0;
    if (fabsf(get_phi(p, 0)) > 1.E-4F && (p->sym90 || p->sym180)) p->sym90 = p->sym180 = 0;
    if (fabsf(g->origin_x) > .01F || fabsf(g->origin_y) > .01F) p->sym90 = p->sym180 = p->symseg = p->syms = 0;
    if (g->min_segment_num > g->max_segment_num) { free(p); return NULL; }
    if (fabsf(get_tantheta(p, 0, 0)) > 1.E-4F) { snprintf(err, errlen, "segment 0 is not direct"); free(p); return NULL; }
  }
  if (p->ntof > 1) {
    /* TOF keeps only the z-shift symmetry (DataSymmetriesForBins_PET_CartesianGrid.cxx:345-366);
     * the flags without that rule are those of the non-TOF sensitivity projector */
    p->nt_sym[0] = p->sym90; p->nt_sym[1] = p->sym180; p->nt_sym[2] = p->symseg; p->nt_sym[3] = p->syms;
    p->sym90 = p->sym180 = p->symseg = p->syms = 0;
  }
  /* find_relation_between_coordinate_systems -- .cxx:51-125 */
  {
    const float ips = g->vox_z;
    const float ppr_f = p->ring_spacing / ips;
    p->ppr = round_f(ppr_f);
    if (fabsf(ppr_f - p->ppr) > 1.E-2) { snprintf(err, errlen, "z spacing must divide the ring spacing"); free(p); return NULL; }
    for (int s = g->min_segment_num; s <= g->max_segment_num; ++s) {
      const int i = SEGI(p, s);
      const float f = p->ax_samp[i] / ips;
      p->ppa[i] = round_f(f);
      if (fabsf(f - p->ppa[i]) > 1.E-2) { snprintf(err, errlen, "z spacing must divide the axial sampling"); free(p); return NULL; }
      const float delta = p->delta[i];
      p->ap2z[i] = (p->max_z + g->min_z) / 2.F - g->origin_z / ips -
                   (p->ppa[i] * (g->num_axial_poss[i] - 1 + 0) + p->ppr * delta) / 2;
    }
  }
  /* PMRT restrictions -- ProjMatrixByBinUsingRayTracing.cxx:276-277 */
  if (fabsf(g->origin_x) > .05F || fabsf(g->origin_y) > .05F) {
    snprintf(err, errlen, "shifted x/y origin is not supported");
    free(p);
    return NULL;
  }
  /* basic rows (the reference's basic-bin cache) */
  const size_t nb = (size_t)p->nseg * g->num_views * (size_t)(2 * p->tmax + 1);
  p->basic = (row*)calloc(nb, sizeof(row));
  int fail = 0;
#pragma omp parallel for schedule(dynamic) collapse(2) reduction(| : fail)
  for (int si = 0; si < p->nseg; ++si)
    for (int v = 0; v < g->num_views; ++v) {
      const int s = si + g->min_segment_num;
      for (int t = -p->tmax; t <= p->tmax; ++t) {
        int bs = s, bv = v, ba = 0, bt = t;
        find_basic_bin(p, &bs, &bv, &ba, &bt);
        if (bs != s || bv != v || bt != t || ba != 0) continue;
        if (calc_row(p, s, v, 0, t, &p->basic[basic_index(p, s, v, t)]) != 0) fail |= 1;
      }
    }
  if (fail) {
    snprintf(err, errlen, "axial sampling is not an integer multiple of the z voxel size");
    or_plan_destroy(p);
    return NULL;
  }
  if (p->ntof > 1) {
#pragma omp parallel for schedule(dynamic) collapse(2)
    for (int si = 0; si < p->nseg; ++si)
      for (int v = 0; v < g->num_views; ++v)
        for (int t = -p->tmax; t <= p->tmax; ++t) {
          row* r = &p->basic[basic_index(p, si + g->min_segment_num, v, t)];
          if (r->e || r->n) tof_weights(p, si + g->min_segment_num, v, t, r);
        }
  }
  return p;
}

void or_plan_destroy(or_plan* p)
{
  if (!p) return;
  if (p->basic) {
    const size_t nb = (size_t)p->nseg * p->g.num_views * (size_t)(2 * p->tmax + 1);
    for (size_t i = 0; i < nb; ++i) row_free(&p->basic[i]);
    free(p->basic);
  }
  free(p->erf_tab);
  free(p);
}

void or_symmetry_flags(const or_plan* p, int32_t* out)
{
  out[0] = p->sym90; out[1] = p->sym180; out[2] = p->symseg; out[3] = p->syms; out[4] = p->symz;
}

int64_t or_num_bins(const or_plan* p) { return p->num_bins * p->ntof; }

int64_t or_bin_index(const or_plan* p, int seg, int view, int ax, int tang)
{
  const int i = SEGI(p, seg);
  return p->seg_offset[i] + ((int64_t)ax * p->g.num_views + view) * p->g.num_tangential_poss +
         (tang - p->g.min_tangential_pos_num);
}

/* basic row + symmetry op of a bin (ProjMatrixByBin.inl:48-80) */
static const row* row_and_op(const or_plan* p, int seg, int view, int ax, int tang, symop* op)
{
  int bs = seg, bv = view, ba = ax, bt = tang;
  *op = sym_op_for_bin(p, &bs, &bv, &ba, &bt);
  return &p->basic[basic_index(p, bs, bv, bt)];
}

/* weight of element k of a basic row for TOF bin kk (the plain value for non-TOF) */
static float elem_w(const or_plan* p, const row* b, int k, int kk)
{
  return (p->ntof > 1 && kk >= 0) ? b->tw[(size_t)k * p->ntof + kk] : b->e[k].w;
}
/* index of a bin in (TOF) projection data: TOF bin outermost */
static int64_t tbin(const or_plan* p, int seg, int view, int ax, int tang, int kk)
{
  return (int64_t)kk * p->num_bins + or_bin_index(p, seg, view, ax, tang);
}

static int export_row(const row* r, int16_t* z, int16_t* y, int16_t* x, float* w, int cap)
{
  if (r->n > cap) return -r->n;
  for (int k = 0; k < r->n; ++k) { z[k] = r->e[k].z; y[k] = r->e[k].y; x[k] = r->e[k].x; w[k] = r->e[k].w; }
  return r->n;
}

int or_get_row(const or_plan* p, int seg, int view, int ax, int tang,
               int16_t* z, int16_t* y, int16_t* x, float* w, int cap)
{
  symop op;
  const row* b = row_and_op(p, seg, view, ax, tang, &op);
  if (b->n > cap) return -b->n;
  for (int k = 0; k < b->n; ++k) {
    int oz, oy, ox;
    apply_op(&op, b->e[k].z, b->e[k].y, b->e[k].x, &oz, &oy, &ox);
    z[k] = (int16_t)oz; y[k] = (int16_t)oy; x[k] = (int16_t)ox; w[k] = b->e[k].w;
  }
  return b->n;
}

/* TOF row of a bin (ProjMatrixByBin::get_proj_matrix_elems_for_one_bin with
 * apply_tof_kernel on the basic bin, ProjMatrixByBin.inl:48-113,115-181):
 * elements of the basic row whose TOF weight is non-zero, transformed. */
int or_get_row_tof(const or_plan* p, int seg, int view, int ax, int tang, int tof,
                   int16_t* z, int16_t* y, int16_t* x, float* w, int cap)
{
  if (p->ntof <= 1 || tof < 0 || tof >= p->ntof) return 0;
  symop op;
  const row* b = row_and_op(p, seg, view, ax, tang, &op);
  int n = 0;
  for (int k = 0; k < b->n; ++k) if (elem_w(p, b, k, tof) != 0.F) ++n;
  if (n > cap) return -n;
  n = 0;
  for (int k = 0; k < b->n; ++k) {
    const float wt = elem_w(p, b, k, tof);
    if (wt == 0.F) continue;
    int oz, oy, ox;
    apply_op(&op, b->e[k].z, b->e[k].y, b->e[k].x, &oz, &oy, &ox);
    z[n] = (int16_t)oz; y[n] = (int16_t)oy; x[n] = (int16_t)ox; w[n] = wt;
    ++n;
  }
  return n;
}

int or_trace_row(const or_plan* p, int seg, int view, int ax, int tang,
                 int16_t* z, int16_t* y, int16_t* x, float* w, int cap)
{
  row r = {0};
  if (calc_row(p, seg, view, ax, tang, &r) != 0) { row_free(&r); return 0; }
  const int n = export_row(&r, z, y, x, w, cap);
  row_free(&r);
  return n;
}

int or_ray_trace(const float* st, const float* sp, const float* vox, float nc,
                 int16_t* z, int16_t* y, int16_t* x, float* w, int cap)
{
  row r = {0};
  ray_trace(&r, st[0], st[1], st[2], sp[0], sp[1], sp[2], vox[0], vox[1], vox[2], nc);
  const int n = export_row(&r, z, y, x, w, cap);
  row_free(&r);
  return n;
}

void or_symmetry_op(const or_plan* p, int seg, int view, int ax, int tang, int32_t* basic4, int32_t* op3)
{
  int bs = seg, bv = view, ba = ax, bt = tang;
  const symop o = sym_op_for_bin(p, &bs, &bv, &ba, &bt);
  basic4[0] = bs; basic4[1] = bv; basic4[2] = ba; basic4[3] = bt;
  op3[0] = o.inplane; op3[1] = o.zsign; op3[2] = o.zadd;
}

void or_bin_geometry(const or_plan* p, int seg, int view, int ax, int tang, float* out)
{
  out[0] = get_s(p, tang);
  out[1] = get_phi(p, view);
  out[2] = get_tantheta(p, seg, tang);
  out[3] = get_costheta(p, seg, tang);
  out[4] = get_t(p, seg, ax, tang);
}

/* ------------------------------------------------------------------ */
/* subsets                                                             */
/* ------------------------------------------------------------------ */
/* list (view,seg) of a subset in the reference's processing order:
 * basic vs (segment ascending, view = subset, subset+n, ...) then related */
static int subset_vs(const or_plan* p, int subset, int nsub, int* views, int* segs, int cap)
{
  int n = 0;
  for (int s = p->g.min_segment_num; s <= p->g.max_segment_num; ++s)
    for (int v = subset; v < p->g.num_views; v += nsub) {
      int bv = v, bs = s;
      if (find_basic_vs(p, &bv, &bs)) continue;
      int rv[16], rs[16];
      const int nr = related_vs(p, v, s, rv, rs);
      for (int k = 0; k < nr; ++k) {
        if (n < cap) { views[n] = rv[k]; segs[n] = rs[k]; }
        ++n;
      }
    }
  return n;
}

int or_subset_view_segments(const or_plan* p, int subset, int num_subsets,
                            int32_t* views, int32_t* segs, int cap)
{
  int* v = (int*)malloc(sizeof(int) * (size_t)(p->nseg * p->g.num_views + 16));
  int* s = (int*)malloc(sizeof(int) * (size_t)(p->nseg * p->g.num_views + 16));
  const int n = subset_vs(p, subset, num_subsets, v, s, p->nseg * p->g.num_views + 16);
  for (int k = 0; k < n && k < cap; ++k) { views[k] = v[k]; segs[k] = s[k]; }
  free(v); free(s);
  return n;
}

/* subsets_are_approximately_balanced -- ...AndProjData.cxx:490-529: the
 * number of segment-0 viewgrams per subset must be equal (0/1 of each
 * related set counted per basic view). Returns 1 when balanced. */
int or_subsets_balanced(const or_plan* p, int nsub)
{
  int first = -1;
  for (int sub = 0; sub < nsub; ++sub) {
    int cnt = 0;
    for (int v = sub; v < p->g.num_views; v += nsub) {
      int bv = v, bs = 0;
      if (find_basic_vs(p, &bv, &bs)) continue;
      int rv[16], rs[16];
      cnt += related_vs(p, v, 0, rv, rs);
    }
    if (first < 0) first = cnt;
    else if (cnt != first) return 0;
  }
  return 1;
}

/* ------------------------------------------------------------------ */
/* projections                                                         */
/* ------------------------------------------------------------------ */
typedef struct { int seg, view; } vs_t;

static int all_vs(const or_plan* p, int subset, int nsub, vs_t** out)
{
  const int cap = p->nseg * p->g.num_views + 16;
  int* v = (int*)malloc(sizeof(int) * (size_t)cap);
  int* s = (int*)malloc(sizeof(int) * (size_t)cap);
  int n;
  if (subset < 0) {
    n = subset_vs(p, 0, 1, v, s, cap);
  } else {
    n = subset_vs(p, subset, nsub, v, s, cap);
  }
  *out = (vs_t*)malloc(sizeof(vs_t) * (size_t)(n > 0 ? n : 1));
  for (int k = 0; k < n; ++k) { (*out)[k].seg = s[k]; (*out)[k].view = v[k]; }
  free(v); free(s);
  return n;
}

static size_t img_index(const or_plan* p, int z, int y, int x)
{
  return ((size_t)(z - p->g.min_z) * p->g.ny + (y - p->g.min_y)) * p->g.nx + (x - p->g.min_x);
}

/* one bin FP: ProjMatrixElemsForOneBin::forward_project -- .cxx:354-371 */
static float fp_bin(const or_plan* p, const float* img, int seg, int view, int ax, int tang, int kk)
{
  symop op;
  const row* b = row_and_op(p, seg, view, ax, tang, &op);
  float acc = 0.f;
  for (int k = 0; k < b->n; ++k) {
    int z, y, x;
    apply_op(&op, b->e[k].z, b->e[k].y, b->e[k].x, &z, &y, &x);
    if (z >= p->g.min_z && z <= p->max_z) acc += img[img_index(p, z, y, x)] * elem_w(p, b, k, kk);
  }
  return acc;
}
static double fp_bin64(const or_plan* p, const double* img, int seg, int view, int ax, int tang, int kk)
{
  symop op;
  const row* b = row_and_op(p, seg, view, ax, tang, &op);
  double acc = 0.;
  for (int k = 0; k < b->n; ++k) {
    int z, y, x;
    apply_op(&op, b->e[k].z, b->e[k].y, b->e[k].x, &z, &y, &x);
    if (z >= p->g.min_z && z <= p->max_z) acc += img[img_index(p, z, y, x)] * (double)elem_w(p, b, k, kk);
  }
  return acc;
}

void or_forward_project(const or_plan* p, const float* img, float* sino, int subset, int nsub)
{
  vs_t* vs; const int n = all_vs(p, subset, nsub, &vs);
  for (int k = 0; k < n; ++k) {
    const int seg = vs[k].seg, view = vs[k].view;
    const int nax = p->g.num_axial_poss[SEGI(p, seg)];
    for (int t = p->g.min_tangential_pos_num; t <= p->max_tang; ++t)
      for (int a = 0; a < nax; ++a)
        for (int kk = 0; kk < p->ntof; ++kk) sino[tbin(p, seg, view, a, t, kk)] = fp_bin(p, img, seg, view, a, t, kk);
  }
  free(vs);
}

void or_forward_project_omp(const or_plan* p, const float* img, float* sino, int subset, int nsub)
{
  vs_t* vs; const int n = all_vs(p, subset, nsub, &vs);
#pragma omp parallel for schedule(dynamic)
  for (int k = 0; k < n; ++k) {
    const int seg = vs[k].seg, view = vs[k].view;
    const int nax = p->g.num_axial_poss[SEGI(p, seg)];
    for (int t = p->g.min_tangential_pos_num; t <= p->max_tang; ++t)
      for (int a = 0; a < nax; ++a)
        for (int kk = 0; kk < p->ntof; ++kk) sino[tbin(p, seg, view, a, t, kk)] = fp_bin(p, img, seg, view, a, t, kk);
  }
  free(vs);
}

void or_forward_project_f64(const or_plan* p, const double* img, double* sino, int subset, int nsub)
{
  vs_t* vs; const int n = all_vs(p, subset, nsub, &vs);
#pragma omp parallel for schedule(dynamic)
  for (int k = 0; k < n; ++k) {
    const int seg = vs[k].seg, view = vs[k].view;
    const int nax = p->g.num_axial_poss[SEGI(p, seg)];
    for (int t = p->g.min_tangential_pos_num; t <= p->max_tang; ++t)
      for (int a = 0; a < nax; ++a)
        for (int kk = 0; kk < p->ntof; ++kk) sino[tbin(p, seg, view, a, t, kk)] = fp_bin64(p, img, seg, view, a, t, kk);
  }
  free(vs);
}

/* ProjMatrixElemsForOneBin::back_project -- .cxx:333-352 (skips zero data) */
static void bp_bin(const or_plan* p, float* img, float data, int seg, int view, int ax, int tang, int kk)
{
  if (data == 0) return;
  symop op;
  const row* b = row_and_op(p, seg, view, ax, tang, &op);
  for (int k = 0; k < b->n; ++k) {
    int z, y, x;
    apply_op(&op, b->e[k].z, b->e[k].y, b->e[k].x, &z, &y, &x);
    if (z >= p->g.min_z && z <= p->max_z) img[img_index(p, z, y, x)] += elem_w(p, b, k, kk) * data;
  }
}
static void bp_bin64(const or_plan* p, double* img, double data, int seg, int view, int ax, int tang, int kk)
{
  if (data == 0) return;
  symop op;
  const row* b = row_and_op(p, seg, view, ax, tang, &op);
  for (int k = 0; k < b->n; ++k) {
    int z, y, x;
    apply_op(&op, b->e[k].z, b->e[k].y, b->e[k].x, &z, &y, &x);
    if (z >= p->g.min_z && z <= p->max_z) img[img_index(p, z, y, x)] += (double)elem_w(p, b, k, kk) * data;
  }
}

/* serial back projection in the reference's bin order
 * (BackProjectorByBinUsingProjMatrixByBin.cxx:107-144: tang outer, axial inner) */
void or_back_project(const or_plan* p, const float* sino, float* img, int subset, int nsub)
{
  memset(img, 0, sizeof(float) * (size_t)p->g.nz * p->g.ny * p->g.nx);
  vs_t* vs; const int n = all_vs(p, subset, nsub, &vs);
  for (int k = 0; k < n; ++k) {
    const int seg = vs[k].seg, view = vs[k].view;
    const int nax = p->g.num_axial_poss[SEGI(p, seg)];
    for (int t = p->g.min_tangential_pos_num; t <= p->max_tang; ++t)
      for (int a = 0; a < nax; ++a)
        for (int kk = 0; kk < p->ntof; ++kk) bp_bin(p, img, sino[tbin(p, seg, view, a, t, kk)], seg, view, a, t, kk);
  }
  free(vs);
}

void or_back_project_f64(const or_plan* p, const double* sino, double* img, int subset, int nsub)
{
  const size_t nv = (size_t)p->g.nz * p->g.ny * p->g.nx;
  memset(img, 0, sizeof(double) * nv);
  vs_t* vs; const int n = all_vs(p, subset, nsub, &vs);
  int nt = 1;
#ifdef _OPENMP
  nt = omp_get_max_threads();
#endif
  double* part = (double*)calloc((size_t)nt * nv, sizeof(double));
#pragma omp parallel for schedule(dynamic)
  for (int k = 0; k < n; ++k) {
    int tid = 0;
#ifdef _OPENMP
    tid = omp_get_thread_num();
#endif
    const int seg = vs[k].seg, view = vs[k].view;
    const int nax = p->g.num_axial_poss[SEGI(p, seg)];
    for (int t = p->g.min_tangential_pos_num; t <= p->max_tang; ++t)
      for (int a = 0; a < nax; ++a)
        for (int kk = 0; kk < p->ntof; ++kk)
          bp_bin64(p, part + (size_t)tid * nv, sino[tbin(p, seg, view, a, t, kk)], seg, view, a, t, kk);
  }
  for (int t = 0; t < nt; ++t)
    for (size_t i = 0; i < nv; ++i) img[i] += part[(size_t)t * nv + i];
  free(part);
  free(vs);
}

/* OpenMP back projection: per-thread images reduced serially
 * (BackProjectorByBin.cxx:316-346) -- the timed CPU baseline */
void or_back_project_omp(const or_plan* p, const float* sino, float* img, int subset, int nsub)
{
  const size_t nv = (size_t)p->g.nz * p->g.ny * p->g.nx;
  memset(img, 0, sizeof(float) * nv);
  vs_t* vs; const int n = all_vs(p, subset, nsub, &vs);
  int nt = 1;
#ifdef _OPENMP
  nt = omp_get_max_threads();
#endif
  float* part = (float*)calloc((size_t)nt * nv, sizeof(float));
#pragma omp parallel for schedule(dynamic)
  for (int k = 0; k < n; ++k) {
    int tid = 0;
#ifdef _OPENMP
    tid = omp_get_thread_num();
#endif
    const int seg = vs[k].seg, view = vs[k].view;
    const int nax = p->g.num_axial_poss[SEGI(p, seg)];
    for (int t = p->g.min_tangential_pos_num; t <= p->max_tang; ++t)
      for (int a = 0; a < nax; ++a)
        for (int kk = 0; kk < p->ntof; ++kk)
          bp_bin(p, part + (size_t)tid * nv, sino[tbin(p, seg, view, a, t, kk)], seg, view, a, t, kk);
  }
  for (int t = 0; t < nt; ++t)
    for (size_t i = 0; i < nv; ++i) img[i] += part[(size_t)t * nv + i];
  free(part);
  free(vs);
}

int64_t or_nnz(const or_plan* p, int subset, int nsub)
{
  vs_t* vs; const int n = all_vs(p, subset, nsub, &vs);
  int64_t total = 0;
#pragma omp parallel for schedule(dynamic) reduction(+ : total)
  for (int k = 0; k < n; ++k) {
    const int seg = vs[k].seg, view = vs[k].view;
    const int nax = p->g.num_axial_poss[SEGI(p, seg)];
    for (int t = p->g.min_tangential_pos_num; t <= p->max_tang; ++t)
      for (int a = 0; a < nax; ++a) {
        symop op;
        const row* b = row_and_op(p, seg, view, a, t, &op);
        for (int e = 0; e < b->n; ++e) {
          const int z = op.zsign * b->e[e].z + op.zadd;
          if (z < p->g.min_z || z > p->max_z) continue;
          if (p->ntof == 1) total++;
          else
            for (int kk = 0; kk < p->ntof; ++kk) total += elem_w(p, b, e, kk) != 0.F;
        }
      }
  }
  free(vs);
  return total;
}

/* ------------------------------------------------------------------ */
/* objective function and OSEM                                          */
/* ------------------------------------------------------------------ */
static int is_end_plane(const or_plan* p, int seg, int ax)
{
  return seg == 0 && (ax == 0 || ax == p->g.num_axial_poss[SEGI(p, 0)] - 1);
}

void or_set_use_tof_sensitivities(or_plan* p, int use) { p->tofsens = use != 0; }

/* Multiplicative factors m of the model ybar = m (A x + b): the normalisation's
 * undo() applied to ones (distributable.cxx:197-215); NULL = none. The array
 * (all TOF bins, measured order) stays owned by the caller. */
void or_set_multiplicative(or_plan* p, const float* mult) { p->mult = mult; }
/* "use subset sensitivities" (...ForMean.cxx:48,374-455) */
void or_set_use_subset_sensitivities(or_plan* p, int use) { p->no_subset_sens = use == 0; }
/* OSMAPOSL keys: "MAP_model" (0 additive, 1 multiplicative), "minimum/maximum
 * relative change" (OSMAPOSLReconstruction.cxx:95-121) */
void or_set_osem_options(or_plan* p, int map_multiplicative, float min_rel, float max_rel)
{
  p->map_multiplicative = map_multiplicative != 0;
  p->min_rel = min_rel;
  p->max_rel = max_rel;
}
/* Prior of OSMAPOSL's one-step-late update: type 0 none, 1 QuadraticPrior,
 * 2 RelativeDifferencePrior; w27 NULL = default weights; kappa caller-owned. */
void or_set_prior(or_plan* p, int type, float beta, float gamma, float eps, const float* kappa, const float* w27,
                  int only_2D)
{
  p->prior_type = type;
  p->beta = beta;
  p->gamma = gamma;
  p->eps = eps;
  p->kappa = kappa;
  p->only_2D = only_2D;
  if (w27) memcpy(p->w27, w27, sizeof(p->w27));
  else or_quadratic_prior_weights(p->g.vox_z, p->g.vox_y, p->g.vox_x, p->w27);
}

/* The multiplicative factor of TOF bin index i: mult viewgrams (1 without a
 * normalisation) with the seg-0 end planes zeroed on request
 * (distributable.cxx:197-223). */
static float mult_at(const or_plan* p, int64_t i, int seg, int ax, int zero_end_planes)
{
  if (zero_end_planes && is_end_plane(p, seg, ax)) return 0.F;
  return p->mult ? p->mult[i] : 1.F;
}

/* With TOF data and "use time-of-flight sensitivities" off (the default,
 * ...AndProjData.cxx:99,533-536,648-670) the sensitivity is the non-TOF back
 * projection of ones: every element with its non-TOF value. */
static void subset_sensitivity_nontof(const or_plan* p, int subset, int nsub, int zero_end_planes, float* img)
{
  memset(img, 0, sizeof(float) * (size_t)p->g.nz * p->g.ny * p->g.nx);
  /* the subset's view-segments are those of the non-TOF projector's symmetries
   * (add_subset_sensitivity uses sens_symmetries_sptr, ...AndProjData.cxx:884-903) */
  or_plan q = *p;
  q.sym90 = p->nt_sym[0]; q.sym180 = p->nt_sym[1]; q.symseg = p->nt_sym[2]; q.syms = p->nt_sym[3];
  vs_t* vs; const int n = all_vs(&q, subset, nsub, &vs);
  for (int k = 0; k < n; ++k) {
    const int seg = vs[k].seg, view = vs[k].view;
    const int nax = p->g.num_axial_poss[SEGI(p, seg)];
    for (int t = p->g.min_tangential_pos_num; t <= p->max_tang; ++t)
      for (int a = 0; a < nax; ++a)
        bp_bin(p, img, mult_at(p, tbin(p, seg, view, a, t, 0), seg, a, zero_end_planes), seg, view, a, t, -1);
  }
  free(vs);
}

void or_subset_sensitivity(const or_plan* p, int subset, int nsub, int zero_end_planes, float* img)
{
  if (p->ntof > 1 && !p->tofsens) {
    subset_sensitivity_nontof(p, subset, nsub, zero_end_planes, img);
    return;
  }
  /* back projection of the mult viewgrams (RPC ...sensitivity_computation,
   * ...AndProjData.cxx:1517-1539) */
  const int64_t nb = p->num_bins * p->ntof;
  float* ones = (float*)malloc(sizeof(float) * (size_t)nb);
  for (int64_t i = 0; i < nb; ++i) ones[i] = p->mult ? p->mult[i] : 1.F;
  if (zero_end_planes) {
    for (int kk = 0; kk < p->ntof; ++kk)
      for (int v = 0; v < p->g.num_views; ++v)
        for (int t = p->g.min_tangential_pos_num; t <= p->max_tang; ++t) {
          ones[tbin(p, 0, v, 0, t, kk)] = 0;
          ones[tbin(p, 0, v, p->g.num_axial_poss[SEGI(p, 0)] - 1, t, kk)] = 0;
        }
  }
  or_back_project(p, ones, img, subset, nsub);
  free(ones);
}

/* divide_and_truncate for one viewgram (segment, view, TOF bin) --
 * recon_array_functions.cxx:172-266 (rim_truncation_sino = 0). Writes the
 * ratio into q. */
static void divide_and_truncate_vg(const or_plan* p, const float* y, const float* ybar, float* q,
                                   int seg, int view, int kk, int zero_end_planes)
{
  const int nax = p->g.num_axial_poss[SEGI(p, seg)];
  float mx = -FLT_MAX;
  for (int a = 0; a < nax; ++a)
    for (int t = p->g.min_tangential_pos_num; t <= p->max_tang; ++t) {
      float yv = y[tbin(p, seg, view, a, t, kk)];
      if (zero_end_planes && is_end_plane(p, seg, a)) yv = 0;
      if (yv > mx) mx = yv;
    }
  const float small_value = maxf_(mx * 0.000001F, 0.F);
  for (int a = 0; a < nax; ++a)
    for (int t = p->g.min_tangential_pos_num; t <= p->max_tang; ++t) {
      const int64_t i = tbin(p, seg, view, a, t, kk);
      float num = y[i];
      if (zero_end_planes && is_end_plane(p, seg, a)) num = 0;
      if (num <= small_value) num = 0;
      else {
        const float max_quotient = 10000.F;
        const float denom = ybar[i];
        if (num > max_quotient * denom) num = max_quotient;
        else num = num / denom;
      }
      q[i] = num;
    }
}

/* ybar += b over one viewgram (b zeroed at end planes when requested) */
static void add_background_vg(const or_plan* p, float* ybar, const float* bkg, int seg, int view, int kk,
                              int zero_end_planes)
{
  const int nax = p->g.num_axial_poss[SEGI(p, seg)];
  for (int a = 0; a < nax; ++a)
    for (int t = p->g.min_tangential_pos_num; t <= p->max_tang; ++t) {
      const int64_t i = tbin(p, seg, view, a, t, kk);
      ybar[i] += (zero_end_planes && is_end_plane(p, seg, a)) ? 0.F : bkg[i];
    }
}

/* The projection data RPC_process_related_viewgrams_gradient back-projects
 * (...AndProjData.cxx:1423-1481): ybar = FP(x) + b, then
 * divide_and_truncate(y, ybar) per viewgram, minus the multiplicative
 * factors (or 1) unless add_sensitivity; bins outside the subset stay 0. */
void or_subset_ratio(const or_plan* p, const float* y, const float* bkg, const float* x,
                     int subset, int nsub, int zero_end_planes, int add_sensitivity, float* q)
{
  const int64_t nb = p->num_bins * p->ntof;
  float* ybar = (float*)calloc((size_t)nb, sizeof(float));
  memset(q, 0, sizeof(float) * (size_t)nb);
  or_forward_project_omp(p, x, ybar, subset, nsub);
  vs_t* vs; const int n = all_vs(p, subset, nsub, &vs);
  for (int k = 0; k < n; ++k) {
    const int seg = vs[k].seg, view = vs[k].view;
    const int nax = p->g.num_axial_poss[SEGI(p, seg)];
    for (int kk = 0; kk < p->ntof; ++kk) {
      if (bkg) add_background_vg(p, ybar, bkg, seg, view, kk, zero_end_planes);
      divide_and_truncate_vg(p, y, ybar, q, seg, view, kk, zero_end_planes);
      if (!add_sensitivity)
        for (int a = 0; a < nax; ++a)
          for (int t = p->g.min_tangential_pos_num; t <= p->max_tang; ++t) {
            const int64_t i = tbin(p, seg, view, a, t, kk);
            q[i] -= mult_at(p, i, seg, a, zero_end_planes);
          }
    }
  }
  free(vs);
  free(ybar);
}

/* compute_sub_gradient_without_penalty[_plus_sensitivity]
 * -- ...AndProjData.cxx:694-735, RPC at :1423-1481 */
void or_subset_gradient(const or_plan* p, const float* y, const float* bkg, const float* x,
                        int subset, int nsub, int zero_end_planes, int add_sensitivity, float* grad)
{
  const int64_t nb = p->num_bins * p->ntof;
  float* q = (float*)malloc(sizeof(float) * (size_t)nb);
  or_subset_ratio(p, y, bkg, x, subset, nsub, zero_end_planes, add_sensitivity, q);
  or_back_project(p, q, grad, subset, nsub);
  free(q);
}

/* divide_and_truncate of numerator / denominator over one viewgram
 * (recon_array_functions.cxx:172-266, rim_truncation_sino = 0): the small
 * value is 1e-6 times the numerator viewgram's maximum (not below 0). */
static void divide_and_truncate_vg2(const or_plan* p, const float* num_in, const float* den, float* q, int seg,
                                    int view, int kk)
{
  const int nax = p->g.num_axial_poss[SEGI(p, seg)];
  float mx = -FLT_MAX;
  for (int a = 0; a < nax; ++a)
    for (int t = p->g.min_tangential_pos_num; t <= p->max_tang; ++t) {
      const float v = num_in[tbin(p, seg, view, a, t, kk)];
      if (v > mx) mx = v;
    }
  const float small_num = maxf_(mx * 0.000001F, 0.F);
  for (int a = 0; a < nax; ++a)
    for (int t = p->g.min_tangential_pos_num; t <= p->max_tang; ++t) {
      const int64_t bi = tbin(p, seg, view, a, t, kk);
      const float nv = num_in[bi];
      const float max_q = 10000.F;
      q[bi] = nv <= small_num ? 0.F : (nv > max_q * den[bi] ? max_q : nv / den[bi]);
    }
}

/* The data-term Hessian products of one subset, accumulated as the reference
 * does (output -= BP(q)), over the subset's viewgrams (all TOF bins):
 *   exact (accumulate_sub_Hessian_times_input_without_penalty,
 *   ...AndProjData.cxx:1066-1218): ybar = FP(current) + b,
 *     q = divide_and_truncate(y * FP(input), ybar * ybar);
 *   approximate (actual_add_multiplication_with_approximate_sub_Hessian_
 *   without_penalty, :957-1064): the data normalised twice (BinNormalisation::
 *   apply divides by the bin efficiency, at least 1e-20, BinNormalisation.cxx:
 *   92-105; the multiplicative factors are those efficiencies),
 *     q = divide_and_truncate(FP(input), y / m / m).
 * The data are used as stored (no end-plane zeroing: the reference reads them
 * with get_related_viewgrams) and no multiplicative factor enters the exact
 * form. Returns 1, with out untouched, when FP(input) has a negative value
 * (the reference aborts then, :1036-1050,1147-1151). */
int or_hessian_times_input(const or_plan* p, const float* y, const float* bkg, const float* current,
                           const float* input, int subset, int nsub, int approximate, float* out)
{
  const int64_t nb = p->num_bins * p->ntof;
  float* fin = (float*)calloc((size_t)nb, sizeof(float));
  float* num = (float*)calloc((size_t)nb, sizeof(float));
  float* den = (float*)calloc((size_t)nb, sizeof(float));
  float* q = (float*)calloc((size_t)nb, sizeof(float));
  const int64_t nv = (int64_t)p->g.nz * p->g.ny * p->g.nx;
  float* bp = (float*)calloc((size_t)nv, sizeof(float));
  int negative = 0;
  or_forward_project_omp(p, input, fin, subset, nsub);
  for (int64_t i = 0; i < nb; ++i)
    if (fin[i] < 0) negative = 1;
  if (!negative) {
    if (!approximate) or_forward_project_omp(p, current, den, subset, nsub);
    vs_t* vs; const int n = all_vs(p, subset, nsub, &vs);
    for (int k = 0; k < n; ++k) {
      const int seg = vs[k].seg, view = vs[k].view;
      const int nax = p->g.num_axial_poss[SEGI(p, seg)];
      for (int kk = 0; kk < p->ntof; ++kk) {
        for (int a = 0; a < nax; ++a)
          for (int t = p->g.min_tangential_pos_num; t <= p->max_tang; ++t) {
            const int64_t i = tbin(p, seg, view, a, t, kk);
            if (approximate) {
              float d = y[i];
              if (p->mult) {
                const float e = maxf_(1.E-20F, p->mult[i]);
                d = d / e;
                d = d / e;
              }
              num[i] = fin[i];
              den[i] = d;
            } else {
              float ybar = den[i];
              if (bkg) ybar = ybar + bkg[i];
              num[i] = y[i] * fin[i];
              den[i] = ybar * ybar;
            }
          }
        divide_and_truncate_vg2(p, num, den, q, seg, view, kk);
      }
    }
    free(vs);
    or_back_project(p, q, bp, subset, nsub);
    for (int64_t j = 0; j < nv; ++j) out[j] = out[j] - bp[j];
  }
  free(fin); free(num); free(den); free(q); free(bp);
  return negative;
}

/* accumulate_loglikelihood -- recon_array_functions.cxx:336-377 via
 * RPC_process_related_viewgrams_accumulate_loglikelihood (...AndProjData.cxx:1484-1515) */
double or_log_likelihood(const or_plan* p, const float* y, const float* bkg, const float* x,
                         int subset, int nsub, int zero_end_planes)
{
  const int64_t nb = p->num_bins * p->ntof;
  float* ybar = (float*)calloc((size_t)nb, sizeof(float));
  or_forward_project_omp(p, x, ybar, subset, nsub);
  vs_t* vs; const int n = all_vs(p, subset, nsub, &vs);
  double total = 0;
  for (int k = 0; k < n; ++k) {
    const int seg = vs[k].seg, view = vs[k].view;
    const int nax = p->g.num_axial_poss[SEGI(p, seg)];
    for (int kk = 0; kk < p->ntof; ++kk) {
      float mx = -FLT_MAX;
      for (int a = 0; a < nax; ++a)
        for (int t = p->g.min_tangential_pos_num; t <= p->max_tang; ++t) {
          float yv = y[tbin(p, seg, view, a, t, kk)];
          if (zero_end_planes && is_end_plane(p, seg, a)) yv = 0;
          if (yv > mx) mx = yv;
        }
      const float small_value = maxf_(mx * 0.000001F, 0.F);
      double result = 0;
      for (int a = 0; a < nax; ++a) {
        double sub = 0;
        const int ep = zero_end_planes && is_end_plane(p, seg, a);
        for (int t = p->g.min_tangential_pos_num; t <= p->max_tang; ++t) {
          const int64_t i = tbin(p, seg, view, a, t, kk);
          float est = ybar[i] + (bkg ? (ep ? 0.F : bkg[i]) : 0.F);
          if (p->mult || zero_end_planes) est = est * mult_at(p, i, seg, a, zero_end_planes);
          const float yv = ep ? 0.F : y[i];
          const float ne = maxf_(est, yv / 10000.F);
          if (yv <= small_value) sub += -(double)ne;
          else sub += yv * log((double)ne) - (double)ne;
        }
        result += sub;
      }
      total += result;
    }
  }
  free(vs);
  free(ybar);
  return total;
}

/* threshold_min_to_small_positive_value -- thresholding.h:101-109 */
static void threshold_min_to_small_positive(float* x, size_t n, float small_number)
{
  float minpos = 0; int found = 0;
  for (size_t i = 0; i < n; ++i)
    if (x[i] > 0 && (!found || x[i] < minpos)) { minpos = x[i]; found = 1; }
  if (found) {
    const float thr = minpos * small_number;
    for (size_t i = 0; i < n; ++i) if (thr > x[i]) x[i] = thr;
  } else {
    for (size_t i = 0; i < n; ++i) x[i] = small_number;
  }
}

/* One-step-late denominator and division of OSMAPOSL with a prior
 * (OSMAPOSLReconstruction.cxx:403-463): d = grad/num_subsets + sens clamped to
 * [sens/10, sens*10] ("additive" MAP model) or sens * clamp(1 + grad, 0.1, 10)
 * ("multiplicative"), grad the prior's gradient at the current estimate;
 * then divide(u, d, 1e-6) (divide.inl:26-47). */
static void osl_divide(const or_plan* p, const float* x, const float* sens, int nsub, float* u, size_t nv)
{
  float* grad = (float*)malloc(sizeof(float) * nv);
  if (p->prior_type == 1)
    or_quadratic_prior(x, p->kappa, p->g.nz, p->g.ny, p->g.nx, p->w27, p->only_2D, p->beta, grad);
  else
    or_relative_difference_prior(x, p->kappa, p->g.nz, p->g.ny, p->g.nx, p->w27, p->only_2D, p->beta, p->gamma,
                                 p->eps, grad);
  for (size_t i = 0; i < nv; ++i) {
    float d = grad[i];
    /* std::max(std::min(d, hi), lo) */
    if (!p->map_multiplicative) {
      d = d / nsub + sens[i];
      const float hi = sens[i] * 10, lo = sens[i] / 10;
      d = (hi < d) ? hi : d;
      d = (d < lo) ? lo : d;
    } else {
      d += 1;
      d = (10.F < d) ? 10.F : d;
      d = (d < 1 / 10.F) ? 1 / 10.F : d;
      d *= sens[i];
    }
    grad[i] = d;
  }
  float mx = -FLT_MAX;
  for (size_t i = 0; i < nv; ++i) if (u[i] > mx) mx = u[i];
  float small_value = mx * 0.000001F;
  small_value = small_value > 0 ? small_value : 0;
  for (size_t i = 0; i < nv; ++i) {
    if (fabsf(grad[i]) <= small_value && fabsf(u[i]) <= small_value) u[i] = 0;
    else u[i] /= grad[i];
  }
  free(grad);
}

void or_osem(const or_plan* p, const float* y, const float* bkg, float* x, int nsub, int nsubiter,
             int start_subset, int zero_end_planes, int use_omp)
{
  const size_t nv = (size_t)p->g.nz * p->g.ny * p->g.nx;
  float** sens = (float**)malloc(sizeof(float*) * (size_t)nsub);
  for (int s = 0; s < nsub; ++s) {
    sens[s] = (float*)malloc(sizeof(float) * nv);
    or_subset_sensitivity(p, s, nsub, zero_end_planes, sens[s]);
  }
  if (p->no_subset_sens) {
    /* "use subset sensitivities" := 0 (...ForMean.cxx:374-455): the total,
     * accumulated subset by subset in float, divided by num_subsets */
    for (int s = 1; s < nsub; ++s)
      for (size_t i = 0; i < nv; ++i) sens[0][i] += sens[s][i];
    for (size_t i = 0; i < nv; ++i) sens[0][i] /= nsub;
    for (int s = 1; s < nsub; ++s) memcpy(sens[s], sens[0], sizeof(float) * nv);
  }
  const int with_prior = p->prior_type != 0 && p->beta != 0;
  float* u = (float*)malloc(sizeof(float) * nv);
  threshold_min_to_small_positive(x, nv, 0.000001F);
  for (int k = 1; k <= nsubiter; ++k) {
    const int subset = (k + start_subset - 1) % nsub;
    if (use_omp) {
      /* same maths, OpenMP back projection */
      const int64_t nb = p->num_bins * p->ntof;
      float* ybar = (float*)calloc((size_t)nb, sizeof(float));
      float* q = (float*)calloc((size_t)nb, sizeof(float));
      or_forward_project_omp(p, x, ybar, subset, nsub);
      vs_t* vs; const int n = all_vs(p, subset, nsub, &vs);
      for (int j = 0; j < n; ++j) {
        const int seg = vs[j].seg, view = vs[j].view;
        for (int kk = 0; kk < p->ntof; ++kk) {
          if (bkg) add_background_vg(p, ybar, bkg, seg, view, kk, zero_end_planes);
          divide_and_truncate_vg(p, y, ybar, q, seg, view, kk, zero_end_planes);
        }
      }
      free(vs);
      or_back_project_omp(p, q, u, subset, nsub);
      free(ybar); free(q);
    } else {
      or_subset_gradient(p, y, bkg, x, subset, nsub, zero_end_planes, 1, u);
    }
    if (with_prior) {
      osl_divide(p, x, sens[subset], nsub, u, nv);
    } else {
      /* divide(update, subsens, 0.F) -- numerics/divide.inl:26-47 */
      float mx = -FLT_MAX;
      for (size_t i = 0; i < nv; ++i) if (u[i] > mx) mx = u[i];
      float small_value = mx * 0.F;
      small_value = small_value > 0 ? small_value : 0;
      for (size_t i = 0; i < nv; ++i) {
        if (fabsf(sens[subset][i]) <= small_value && fabsf(u[i]) <= small_value) u[i] = 0;
        else u[i] /= sens[subset][i];
      }
    }
    /* threshold_upper_lower(min_rel, max_rel) except subiteration 1
     * (OSMAPOSLReconstruction.cxx:491-507, thresholding.h:45-54) */
    if (k != 1)
      for (size_t i = 0; i < nv; ++i) {
        if (u[i] > p->max_rel) u[i] = p->max_rel;
        else if (p->min_rel > u[i]) u[i] = p->min_rel;
      }
    for (size_t i = 0; i < nv; ++i) x[i] *= u[i];
  }
  free(u);
  for (int s = 0; s < nsub; ++s) free(sens[s]);
  free(sens);
}

/* ------------------------------------------------------------------ */
/* QuadraticPrior (C2 loss term)                                       */
/* ------------------------------------------------------------------ */

/* compute_weights (QuadraticPrior.cxx:202-226): w = vox_x / |offset|, 0 at
 * the centre; weights[dz+1][dy+1][dx+1]. The squared length is summed in
 * float; the unqualified sqrt there takes the C library's double overload
 * (no <math.h> reaches that file, only <cmath>), so the quotient is formed in
 * double and rounded to float once. The float overload would differ in the
 * last bit for some offsets (parity unpinned: a build of the reference would
 * settle it). */
void or_quadratic_prior_weights(float vox_z, float vox_y, float vox_x, float* w27) {
  for (int z = -1; z <= 1; ++z)
    for (int y = -1; y <= 1; ++y)
      for (int x = -1; x <= 1; ++x) {
        float* w = &w27[((z + 1) * 3 + (y + 1)) * 3 + (x + 1)];
        if (z == 0 && y == 0 && x == 0) {
          *w = 0;
        } else {
          const float sx = (float)x * vox_x, sy = (float)y * vox_y, sz = (float)z * vox_z;
          *w = (float)((double)vox_x / sqrt((double)(sx * sx + sy * sy + sz * sz)));
        }
      }
}

/* compute_value (QuadraticPrior.cxx:229-298) and compute_gradient
 * (:301-367) over an image [nz][ny][nx]; neighbours outside the image are
 * skipped; kappa may be NULL. Per term: float w*(a-b)^2/4 (value) or
 * float w*(a-b) (gradient), accumulated in double; gradient scaled by beta
 * in double then rounded to float. */
double or_quadratic_prior(const float* img, const float* kappa, int nz, int ny, int nx, const float* w27,
                          int only_2D, float beta, float* grad) {
  double result = 0.;
  const int wz0 = only_2D ? 0 : -1, wz1 = only_2D ? 0 : 1;
  for (int z = 0; z < nz; ++z) {
    const int min_dz = wz0 > -z ? wz0 : -z, max_dz = wz1 < nz - 1 - z ? wz1 : nz - 1 - z;
    for (int y = 0; y < ny; ++y) {
      const int min_dy = -1 > -y ? -1 : -y, max_dy = 1 < ny - 1 - y ? 1 : ny - 1 - y;
      for (int x = 0; x < nx; ++x) {
        const int min_dx = -1 > -x ? -1 : -x, max_dx = 1 < nx - 1 - x ? 1 : nx - 1 - x;
        const int64_t v = ((int64_t)z * ny + y) * nx + x;
        double g = 0.;
        for (int dz = min_dz; dz <= max_dz; ++dz)
          for (int dy = min_dy; dy <= max_dy; ++dy)
            for (int dx = min_dx; dx <= max_dx; ++dx) {
              const int64_t n = ((int64_t)(z + dz) * ny + (y + dy)) * nx + (x + dx);
              const float w = w27[((dz + 1) * 3 + (dy + 1)) * 3 + (dx + 1)];
              const float d = img[v] - img[n];
              double cv = (double)(w * (d * d) / 4);
              double cg = (double)(w * d);
              if (kappa) {
                const float kk = kappa[v] * kappa[n];
                cv *= kk;
                cg *= kk;
              }
              result += cv;
              g += cg;
            }
        if (grad) grad[v] = (float)(g * beta);
      }
    }
  }
  return result * beta;
}

/* ------------------------------------------------------------------ */
/* RelativeDifferencePrior                                              */
/* ------------------------------------------------------------------ */
/* compute_value (RelativeDifferencePrior.cxx:303-373) and compute_gradient
 * (:375-433) over an image [nz][ny][nx] with the weights of compute_weights
 * (:250-278, the same as the quadratic prior's); neighbours outside the image
 * are skipped; kappa may be NULL.
 *   value(x, y) = 0.5 * (double)(x-y)^2 / (float)(x + y + gamma|x-y| + eps)  (:280-284)
 *   derivative_10(x, y) = (float)((double)(x-y) * (float)(gamma|x-y| + x + 3y + 2eps)
 *                         / ((double)(x+y) + gamma|x-y| + eps)^2), 0 at x = y = 0, eps = 0  (:286-300)
 * Terms are summed in double in (dz, dy, dx) order per voxel; the value's
 * terms at x = y = 0 (eps = 0) are 0. Returns beta * value; grad (if not NULL)
 * = (float)(beta * gradient). */
double or_relative_difference_prior(const float* img, const float* kappa, int nz, int ny, int nx, const float* w27,
                                    int only_2D, float beta, float gamma, float eps, float* grad)
{
  double result = 0.;
  const int wz0 = only_2D ? 0 : -1, wz1 = only_2D ? 0 : 1;
  for (int z = 0; z < nz; ++z) {
    const int min_dz = wz0 > -z ? wz0 : -z, max_dz = wz1 < nz - 1 - z ? wz1 : nz - 1 - z;
    for (int y = 0; y < ny; ++y) {
      const int min_dy = -1 > -y ? -1 : -y, max_dy = 1 < ny - 1 - y ? 1 : ny - 1 - y;
      for (int x = 0; x < nx; ++x) {
        const int min_dx = -1 > -x ? -1 : -x, max_dx = 1 < nx - 1 - x ? 1 : nx - 1 - x;
        const int64_t v = ((int64_t)z * ny + y) * nx + x;
        double g = 0.;
        for (int dz = min_dz; dz <= max_dz; ++dz)
          for (int dy = min_dy; dy <= max_dy; ++dy)
            for (int dx = min_dx; dx <= max_dx; ++dx) {
              const int64_t n = ((int64_t)(z + dz) * ny + (y + dy)) * nx + (x + dx);
              const float w = w27[((dz + 1) * 3 + (dy + 1)) * 3 + (dx + 1)];
              const float a = img[v], b = img[n];
              const float ad = fabsf(a - b);
              double cv;
              if (eps == 0.F && a == 0.F && b == 0.F) {
                cv = 0.0;
              } else {
                const float den = a + b + gamma * ad + eps;
                const double dd = (double)(a - b);
                cv = (double)w * (0.5 * (dd * dd / den));
              }
              float der;
              if (eps == 0.F && a == 0.F && b == 0.F) {
                der = 0.F;
              } else {
                const float t = gamma * ad + a + 3 * b + 2 * eps;
                const double num = (double)(a - b) * t;
                const double ds = (double)(a + b) + gamma * ad + eps;
                der = (float)(num / (ds * ds));
              }
              double cg = (double)(w * der);
              if (kappa) {
                const float kk = kappa[v] * kappa[n];
                cv *= kk;
                cg *= kk;
              }
              result += cv;
              g += cg;
            }
        if (grad) grad[v] = (float)(g * beta);
      }
    }
  }
  return result * beta;
}

/* ------------------------------------------------------------------ */
/* List mode (PoissonLogLikelihoodWithLinearModelForMeanAndListModeData */
/* WithProjMatrixByBin)                                                 */
/* ------------------------------------------------------------------ */
/* Whether an event's bin lies in the projection data (the cache keeps only
 * those, ...ListModeDataWithProjMatrixByBin.cxx:500-511). */
static int lm_event_in_range(const or_plan* p, const int32_t* e)
{
  const int seg = e[0], view = e[1], ax = e[2], tang = e[3], tof = e[4];
  if (seg < p->g.min_segment_num || seg > p->g.max_segment_num) return 0;
  if (view < 0 || view >= p->g.num_views) return 0;
  if (tang < p->g.min_tangential_pos_num || tang > p->max_tang) return 0;
  if (ax < 0 || ax >= p->g.num_axial_poss[SEGI(p, seg)]) return 0;
  const int min_tof = -(p->ntof / 2);
  if (p->ntof > 1 ? (tof < min_tof || tof > min_tof + p->ntof - 1) : tof != 0) return 0;
  return 1;
}

/* LM_gradient_and_value over the prompt events of one subset, in event order
 * (distributable.txx:99-140 with ...ListModeDataWithProjMatrixByBin.cxx:748-787):
 * events n x (segment, view, axial, tangential, timing position); an event is
 * in subset s when its basic bin's view % num_subsets == s (num_subsets > 1);
 * fwd = row . x + additive (the additive term of the event's bin; for TOF
 * data the cache keeps that of the last timing position, .cxx:570-596).
 * The reference instantiates the function twice: the gradient with
 * <do_gradient=true, do_value=false> (:822-843), whose early return for
 * 1 > 1e4 fwd sits inside `if (do_value)`, so every event back-projects
 * row^T (1/fwd); the value with <false, true> (:876-911): value -=
 * log(1/1e4) when 1 > 1e4 fwd, else value -= log(fwd). add may be NULL (no
 * additive term). grad (if not NULL) is accumulated into. Returns the value. */
double or_lm_gradient(const or_plan* p, const int32_t* ev, int64_t n, const float* add, const float* x, int subset,
                      int nsub, float* grad)
{
  const float max_quotient = 10000.F;
  double value = 0.;
  for (int64_t i = 0; i < n; ++i) {
    const int32_t* e = ev + 5 * i;
    if (!lm_event_in_range(p, e)) continue;
    const int seg = e[0], view = e[1], ax = e[2], tang = e[3];
    const int kk = p->ntof > 1 ? e[4] + p->ntof / 2 : 0;
    if (nsub > 1) {
      int bs = seg, bv = view, ba = ax, bt = tang;
      (void)sym_op_for_bin(p, &bs, &bv, &ba, &bt);
      if (bv % nsub != subset) continue;
    }
    const float add_term = add ? add[tbin(p, seg, view, ax, tang, p->ntof - 1)] : 0.F;
    const float fwd = fp_bin(p, x, seg, view, ax, tang, p->ntof > 1 ? kk : -1) + add_term;
    const float num = 1.F;
    if (grad) bp_bin(p, grad, num / fwd, seg, view, ax, tang, p->ntof > 1 ? kk : -1);
    if (num > max_quotient * fwd) value -= num * log((double)(num / max_quotient));
    else value -= num * log((double)fwd);
  }
  return value;
}

/* LM_Hessian over the subset's events, accumulated into out
 * (actual_accumulate_sub_Hessian_times_input_without_penalty,
 * ...ListModeDataWithProjMatrixByBin.cxx:790-822,845-863,963-1003): per event
 * (bin value 1), fwd = row . x + the additive term; an event with
 * 1 > 1e4 fwd is skipped; r = row . rhs, skipped when 0; otherwise
 * row^T (-1 / (fwd * fwd) * r) is added to out. */
void or_lm_hessian(const or_plan* p, const int32_t* ev, int64_t n, const float* add, const float* x,
                   const float* rhs, int subset, int nsub, float* out)
{
  const float max_quotient = 10000.F;
  for (int64_t i = 0; i < n; ++i) {
    const int32_t* e = ev + 5 * i;
    if (!lm_event_in_range(p, e)) continue;
    const int seg = e[0], view = e[1], ax = e[2], tang = e[3];
    const int kk = p->ntof > 1 ? e[4] + p->ntof / 2 : -1;
    if (nsub > 1) {
      int bs = seg, bv = view, ba = ax, bt = tang;
      (void)sym_op_for_bin(p, &bs, &bv, &ba, &bt);
      if (bv % nsub != subset) continue;
    }
    const float add_term = add ? add[tbin(p, seg, view, ax, tang, p->ntof - 1)] : 0.F;
    const float fwd = fp_bin(p, x, seg, view, ax, tang, kk) + add_term;
    const float ev_value = 1.F;
    if (ev_value > max_quotient * fwd) continue;
    const float fwd_sq = fwd * fwd;
    const float minus_y_over_fwd2 = -ev_value / fwd_sq;
    const float r = fp_bin(p, rhs, seg, view, ax, tang, kk);
    if (r == 0) continue;
    bp_bin(p, out, minus_y_over_fwd2 * r, seg, view, ax, tang, kk);
  }
}

/* OSMAPOSL on list-mode data, prior-free: subset sensitivities BP(m or 1)
 * without end-plane zeroing (add_subset_sensitivity, .cxx:661-720), update
 * x *= divide(sum over the subset's events of row^T/fwd, sens_s, 0),
 * thresholded except in subiteration 1. */
void or_lm_osem(const or_plan* p, const int32_t* ev, int64_t n, const float* add, float* x, int nsub, int nsubiter,
                int start_subset)
{
  const size_t nv = (size_t)p->g.nz * p->g.ny * p->g.nx;
  float** sens = (float**)malloc(sizeof(float*) * (size_t)nsub);
  for (int s = 0; s < nsub; ++s) {
    sens[s] = (float*)malloc(sizeof(float) * nv);
    or_subset_sensitivity(p, s, nsub, 0, sens[s]);
  }
  float* u = (float*)malloc(sizeof(float) * nv);
  threshold_min_to_small_positive(x, nv, 0.000001F);
  for (int k = 1; k <= nsubiter; ++k) {
    const int subset = (k + start_subset - 1) % nsub;
    memset(u, 0, sizeof(float) * nv);
    (void)or_lm_gradient(p, ev, n, add, x, subset, nsub, u);
    for (size_t i = 0; i < nv; ++i) {
      if (fabsf(sens[subset][i]) <= 0.F && fabsf(u[i]) <= 0.F) u[i] = 0;
      else u[i] /= sens[subset][i];
    }
    if (k != 1)
      for (size_t i = 0; i < nv; ++i) {
        if (u[i] > p->max_rel) u[i] = p->max_rel;
        else if (p->min_rel > u[i]) u[i] = p->min_rel;
      }
    for (size_t i = 0; i < nv; ++i) x[i] *= u[i];
  }
  free(u);
  for (int s = 0; s < nsub; ++s) free(sens[s]);
  free(sens);
}

/* ------------------------------------------------------------------ */
/* Prior Hessian products                                               */
/* ------------------------------------------------------------------ */
/* derivative_20 / derivative_11 of the relative difference prior
 * (RelativeDifferencePrior.cxx:617-641): pow of float arguments is the double
 * pow; a*a of a float is exact in double and (a*a)*a is a^3 correctly rounded. */
static float rdp_d20(float xj, float xk, float gamma, float eps)
{
  if (!(xj > 0.F || xk > 0.F || eps > 0.F)) return INFINITY;
  const double a = (double)(2 * xk + eps);
  const double d = (double)(xj + xk + gamma * fabsf(xj - xk) + eps);
  return (float)(2 * (a * a) / ((d * d) * d));
}
static float rdp_d11(float xj, float xk, float gamma, float eps)
{
  if (!(xj > 0.F || xk > 0.F || eps > 0.F)) return INFINITY;
  const float num = -2 * (2 * xj + eps) * (2 * xk + eps);
  const double d = (double)(xj + xk + gamma * fabsf(xj - xk) + eps);
  return (float)((double)num / ((d * d) * d));
}

/* accumulate_Hessian_times_input (QuadraticPrior.cxx:604-690,
 * RelativeDifferencePrior.cxx:535-614): out[j] += beta * sum over
 * neighbours k with w != 0 of w (f20(x_j, x_k) in_j + f11(x_j, x_k) in_k)
 * kappa_j kappa_k, float arithmetic, (dz, dy, dx) order. kind 1: quadratic
 * (f20 = 1, f11 = -1); kind 2: relative difference. kind 3: the quadratic
 * prior's add_multiplication_with_approximate_Hessian (:537-602): out[j] +=
 * beta * sum w in_k kappa_j kappa_k (the centre included, weight 0). */
void or_prior_hessian_times_input(int kind, const float* x, const float* in, const float* kappa, int nz, int ny, int nx,
                                  const float* w27, int only_2D, float beta, float gamma, float eps, float* out)
{
  const int wz0 = only_2D ? 0 : -1, wz1 = only_2D ? 0 : 1;
  for (int z = 0; z < nz; ++z) {
    const int min_dz = wz0 > -z ? wz0 : -z, max_dz = wz1 < nz - 1 - z ? wz1 : nz - 1 - z;
    for (int y = 0; y < ny; ++y) {
      const int min_dy = -1 > -y ? -1 : -y, max_dy = 1 < ny - 1 - y ? 1 : ny - 1 - y;
      for (int xx = 0; xx < nx; ++xx) {
        const int min_dx = -1 > -xx ? -1 : -xx, max_dx = 1 < nx - 1 - xx ? 1 : nx - 1 - xx;
        const int64_t v = ((int64_t)z * ny + y) * nx + xx;
        float result = 0;
        for (int dz = min_dz; dz <= max_dz; ++dz)
          for (int dy = min_dy; dy <= max_dy; ++dy)
            for (int dx = min_dx; dx <= max_dx; ++dx) {
              const int64_t n = ((int64_t)(z + dz) * ny + (y + dy)) * nx + (xx + dx);
              float current = w27[((dz + 1) * 3 + (dy + 1)) * 3 + (dx + 1)];
              if (kind == 3) {
                current = current * in[n];
              } else {
                if (current == 0.F) continue;
                const float f20 = kind == 1 ? 1.F : rdp_d20(x[v], x[n], gamma, eps);
                const float f11 = kind == 1 ? -1.F : rdp_d11(x[v], x[n], gamma, eps);
                if (dz == 0 && dy == 0 && dx == 0) current *= f20 * in[v];
                else current *= (f20 * in[v] + f11 * in[n]);
              }
              if (kappa) current *= kappa[v] * kappa[n];
              result += current;
            }
        out[v] += result * beta;
      }
    }
  }
}

