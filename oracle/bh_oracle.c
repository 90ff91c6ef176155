/*
 * bh_oracle.c — FP64 CPU oracle for the photon-geodesic march.
 *
 * TEST INFRASTRUCTURE ONLY (results are checked against it).  Loaded by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg; never by the
 * product library.  Built with -O2 -ffp-contract=off so every product and sum
 * rounds exactly once, in the order written, like the reference's scalar code.
 *
 * What it restates (reference paths relative to the reference repo root):
 *   camera          include/blackhole/camera.h:37-63
 *   ray set-up      include/blackhole/blackhole_solution_test.cc:167-211
 *   G, SolveG, b_c  include/blackhole/blackhole_solution.h:24-53
 *   march           include/blackhole/blackhole_solution_test.cc:217-296
 *   FindCollision   include/blackhole/object/object_manager.h:69-85
 *   horizon test    include/blackhole/blackhole_solution.h:65-88
 *   annulus test    include/blackhole/object/vector_object.h:328-347
 *   texel lookup    include/blackhole/object/vector_object.h:163-179
 *   flat march      include/blackhole/ray_tracer.h:17-96,
 *                   include/blackhole/object/vector_object.h:107-239,
 *                   include/blackhole/object/pattern.h:14-50
 * OpenCV arithmetic used by those lines (cv::normalize = v * (1/|v|),
 * Vec / s = Vec * (1/s), Matx33::inv = adjugate * (1/det), dot and matvec
 * summed left to right) is restated from OpenCV's public headers; OpenCV is
 * not in the reference tree and its exact rounding is not pinned by any
 * reference test.
 *
 * Extensions that the reference does not define (documented, shared with
 * the GPU kernels): escape to r >= r_escape with a step cap and an
 * equirectangular background, textured spheres, bilinear filtering, the
 * exact-root mode, the degenerate-ray rule, texel clamping.
 */
#include "bh_oracle.h"

#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#define OR_PI 3.141592653589793238462643383279

/* ------------------------------------------------------------------------ */
/* 3-vectors in OpenCV's cv::Vec3d arithmetic                               */
/* ------------------------------------------------------------------------ */
typedef struct { double v[3]; } v3;

static inline v3 mk(double a, double b, double c) { v3 r = {{a, b, c}}; return r; }
static inline v3 ld(const double* p) { return mk(p[0], p[1], p[2]); }
static inline v3 add(v3 a, v3 b) { return mk(a.v[0] + b.v[0], a.v[1] + b.v[1], a.v[2] + b.v[2]); }
static inline v3 sub(v3 a, v3 b) { return mk(a.v[0] - b.v[0], a.v[1] - b.v[1], a.v[2] - b.v[2]); }
static inline v3 scl(v3 a, double s) { return mk(a.v[0] * s, a.v[1] * s, a.v[2] * s); }
/* Matx::dot: s = 0; s += a_i * b_i, i = 0..2 */
static inline double dot(v3 a, v3 b) {
  double s = 0;
  s += a.v[0] * b.v[0];
  s += a.v[1] * b.v[1];
  s += a.v[2] * b.v[2];
  return s;
}
static inline v3 cross(v3 a, v3 b) {
  return mk(a.v[1] * b.v[2] - a.v[2] * b.v[1], a.v[2] * b.v[0] - a.v[0] * b.v[2],
            a.v[0] * b.v[1] - a.v[1] * b.v[0]);
}
/* cv::normalize(Vec): v * (|v| ? 1/|v| : 0), |v| = sqrt(normL2Sqr) */
static inline v3 normalize(v3 a) {
  double nv = sqrt(dot(a, a));
  return scl(a, nv != 0.0 ? 1. / nv : 0.);
}
/* Matx33 * Vec3: row sums from zero, left to right */
static inline v3 matvec(const double m[9], v3 x) {
  v3 r;
  for (int i = 0; i < 3; ++i) {
    double s = 0;
    s += m[3 * i + 0] * x.v[0];
    s += m[3 * i + 1] * x.v[1];
    s += m[3 * i + 2] * x.v[2];
    r.v[i] = s;
  }
  return r;
}

/* Matx33::inv via determinant and adjugate; singular -> zeros, ok = 0. */
void or_inv3(const double a[9], double b[9], int* ok) {
#define A(i, j) a[3 * (i) + (j)]
#define B(i, j) b[3 * (i) + (j)]
  double d = A(0, 0) * (A(1, 1) * A(2, 2) - A(2, 1) * A(1, 2)) -
             A(0, 1) * (A(1, 0) * A(2, 2) - A(2, 0) * A(1, 2)) +
             A(0, 2) * (A(1, 0) * A(2, 1) - A(2, 0) * A(1, 1));
  if (d == 0) {
    memset(b, 0, 9 * sizeof(double));
    if (ok) *ok = 0;
    return;
  }
  d = 1 / d;
  B(0, 0) = (A(1, 1) * A(2, 2) - A(1, 2) * A(2, 1)) * d;
  B(0, 1) = (A(0, 2) * A(2, 1) - A(0, 1) * A(2, 2)) * d;
  B(0, 2) = (A(0, 1) * A(1, 2) - A(0, 2) * A(1, 1)) * d;
  B(1, 0) = (A(1, 2) * A(2, 0) - A(1, 0) * A(2, 2)) * d;
  B(1, 1) = (A(0, 0) * A(2, 2) - A(0, 2) * A(2, 0)) * d;
  B(1, 2) = (A(0, 2) * A(1, 0) - A(0, 0) * A(1, 2)) * d;
  B(2, 0) = (A(1, 0) * A(2, 1) - A(1, 1) * A(2, 0)) * d;
  B(2, 1) = (A(0, 1) * A(2, 0) - A(0, 0) * A(2, 1)) * d;
  B(2, 2) = (A(0, 0) * A(1, 1) - A(0, 1) * A(1, 0)) * d;
  if (ok) *ok = 1;
#undef A
#undef B
}

/* ------------------------------------------------------------------------ */
/* Camera (camera.h:37-63)                                                   */
/* ------------------------------------------------------------------------ */
double or_focal_length(int width, double fov) {
  double hi = nextafter(OR_PI, 0.0);
  double f = fov;
  if (!(0.0 < f)) f = 0.0; /* std::max((T)0, fov) keeps 0 for NaN too */
  if (hi < f) f = hi;      /* std::min(fov, nextafter(pi, 0))          */
  return width / (2 * tan(f / 2.0));
}

static v3 focus_vector(const bh_camera* c) {
  return scl(ld(c->axis_x), or_focal_length(c->width, c->fov));
}

/* PixelVector(x, y, fv) = fv - v_y (W/2 - x) - v_z (H/2 - y) (camera.h:55-59) */
static v3 pixel_vector(const bh_camera* c, v3 fv, int x, int y) {
  double sx = c->width / 2.0 - x;
  double sy = c->height / 2.0 - y;
  return sub(sub(fv, scl(ld(c->axis_y), sx)), scl(ld(c->axis_z), sy));
}

void or_pixel_vector(const bh_camera* cam, int x, int y, double out[3]) {
  v3 p = pixel_vector(cam, focus_vector(cam), x, y);
  memcpy(out, p.v, sizeof p.v);
}

/* ------------------------------------------------------------------------ */
/* StaticBlackhole (blackhole_solution.h:24-53)                              */
/* ------------------------------------------------------------------------ */
double or_G(double u, double b, double M) {
  return (u * u * (2.0 * M * u - 1.0)) + (1.0 / (b * b));
}

double or_b_c(double M) { return 3.0 * sqrt(3.0) * M; }

/* SolveG: bisection on [cbrt(eps_double), 1/(3M)], lower end returned.
 * root_mode BH_ROOT_EXACT keeps bisecting until the bracket stops shrinking,
 * which makes the turning point a smooth function of (b, M). */
double or_solve_g(double b, double M, int root_mode) {
  double l = 0.0 + cbrt(DBL_EPSILON);
  double r = 1.0 / (3.0 * M);
  int count = root_mode == BH_ROOT_EXACT ? 200 : 20;
  for (int i = 0; i < count; ++i) {
    double mid = (l + r) / 2.0;
    if (root_mode == BH_ROOT_EXACT && (mid == l || mid == r)) break;
    if (or_G(mid, b, M) > 0.0)
      l = mid;
    else
      r = mid;
  }
  return l;
}

/* ------------------------------------------------------------------------ */
/* Ray set-up (blackhole_solution_test.cc:167-211)                           */
/* ------------------------------------------------------------------------ */
int or_ray_setup(const bh_scene* s, const bh_options* o, int x, int y, or_ray* ray) {
  const bh_camera* cam = &s->camera;
  v3 B = ld(s->hole.position);
  double M = s->hole.mass;
  v3 fv = focus_vector(cam);
  v3 pv = sub(pixel_vector(cam, fv, x, y), B);
  v3 F = sub(ld(cam->position), B);

  v3 yv = normalize(sub(F, pv));
  v3 zv = normalize(cross(pv, yv));
  if (isnan(zv.v[0])) zv = mk(1, 0, 0);
  v3 xv = normalize(cross(yv, zv));

  double R[9] = {zv.v[0], yv.v[0], xv.v[0], zv.v[1], yv.v[1], xv.v[1],
                 zv.v[2], yv.v[2], xv.v[2]};
  int ok = 1;
  or_inv3(R, ray->Rinv, &ok);
  memcpy(ray->R, R, sizeof R);
  memcpy(ray->pv, pv.v, sizeof pv.v);
  memcpy(ray->F, F.v, sizeof F.v);
  v3 Fp = matvec(ray->Rinv, F);
  memcpy(ray->Fp, Fp.v, sizeof Fp.v);

  /* Degenerate frame (the pixel line passes through B): the reference
   * marches NaNs; defined here as captured / black. */
  ray->degenerate = !ok || dot(zv, zv) == 0.0;

  double b = Fp.v[2];
  double bc = or_b_c(M);
  double sol = -1;
  if (b >= bc) sol = or_solve_g(b, M, o->root);
  ray->captured = sol < 0;
  ray->u_end = b < bc ? 1.0 / (3 * M) : sol;
  ray->b = b;
  ray->r0 = sqrt(dot(Fp, Fp));
  ray->phi0 = atan(Fp.v[2] / Fp.v[1]);
  ray->u0 = 1. / ray->r0;
  int nstep = o->nstep > 0 ? o->nstep : 20;
  ray->du = (ray->u_end - ray->u0) * (1.0 / nstep);
  ray->du_h = ray->du / 2.;
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Objects                                                                   */
/* ------------------------------------------------------------------------ */
/* StaticBlackhole::Collide (blackhole_solution.h:65-88) for centre c and
 * radius R; the same segment/sphere test is used for textured spheres. */
static int seg_sphere(v3 q1, v3 q2, v3 c, double R, v3* inter) {
  v3 p1 = sub(q1, c);
  v3 p2 = sub(q2, c);
  double d1 = sqrt(dot(p1, p1));
  double d2 = sqrt(dot(p2, p2));
  if (d1 < R && d2 < R) return 0;
  v3 Q1 = sub(q1, c);
  v3 Q2 = sub(q2, c);
  v3 Q = sub(Q2, Q1);
  double qd = dot(Q, Q1);
  double cc = qd * qd - dot(Q, Q) * (dot(Q1, Q1) - R * R);
  if (cc <= 0) return 0;
  double t = (((double)1) / dot(Q, Q)) * (-dot(Q, Q1) - sqrt(cc));
  if (t <= 0 || t >= 1) return 0;
  *inter = add(add(Q1, scl(Q, t)), c);
  return 1;
}

int or_seg_sphere(const double q1[3], const double q2[3], const double c[3], double R,
                  double inter[3]) {
  v3 r;
  int h = seg_sphere(ld(q1), ld(q2), ld(c), R, &r);
  if (h) memcpy(inter, r.v, sizeof r.v);
  return h;
}

/* Annulus::Collide (vector_object.h:328-347): sign change of n.p across the
 * chord, crossing by linear interpolation on the chord, radius window. */
static int annulus_collide(const bh_annulus* a, v3 q1, v3 q2, v3* inter) {
  v3 c0 = ld(a->vertex[0]);
  v3 n = ld(a->normal);
  v3 p1 = sub(q1, c0);
  v3 p2 = sub(q2, c0);
  if (dot(n, p1) * dot(n, p2) >= 0) return 0;
  double t1 = dot(n, p1);
  double t2 = dot(n, p2);
  v3 num = add(scl(p1, fabs(t2)), scl(p2, fabs(t1)));
  v3 c = scl(num, 1. / (fabs(t1) + fabs(t2)));
  if (sqrt(dot(c, c)) > a->r_outer) return 0;
  if (sqrt(dot(c, c)) < a->r_inner) return 0;
  *inter = add(c, c0);
  return 1;
}

int or_annulus_collide(const bh_annulus* a, const double q1[3], const double q2[3],
                       double inter[3]) {
  v3 r;
  int h = annulus_collide(a, ld(q1), ld(q2), &r);
  if (h) memcpy(inter, r.v, sizeof r.v);
  return h;
}

/* Rectangle::color (vector_object.h:163-179): continuous texel coordinates
 * and the truncated (int) indices, clamped into the image (the reference
 * reads out of bounds at the far rim). */
void or_rect_texel(const double vtx[5][3], int rows, int cols, const double pp[3],
                   double* fw, double* fh, int* pw, int* ph) {
  v3 v1 = ld(vtx[1]);
  v3 s1 = sub(ld(vtx[2]), v1);
  v3 s2 = sub(ld(vtx[4]), v1);
  v3 v = sub(ld(pp), v1);
  double r = sqrt(dot(v, v));
  double theta = acos(dot(s1, v) / (sqrt(dot(s1, s1)) * r));
  theta = isnan(theta) ? 0 : theta;
  double a = (r * cos(theta) / sqrt(dot(s1, s1))) * cols;
  double b = (r * sin(theta) / sqrt(dot(s2, s2))) * rows;
  int iw = (int)a, ih = (int)b;
  if (iw < 0) iw = 0;
  if (iw > cols - 1) iw = cols - 1;
  if (ih < 0) ih = 0;
  if (ih > rows - 1) ih = rows - 1;
  if (fw) *fw = a;
  if (fh) *fh = b;
  if (pw) *pw = iw;
  if (ph) *ph = ih;
}

/* Equirectangular map of a unit direction (extension): longitude
 * atan2(d_y, d_x) over [-pi, pi) -> columns, latitude asin(d_z) over
 * [pi/2, -pi/2] -> rows, nearest texel, clamped. */
void or_equirect_texel(const double d[3], int rows, int cols, double* fc, double* fr,
                       int* col, int* row) {
  double z = d[2] < -1.0 ? -1.0 : (d[2] > 1.0 ? 1.0 : d[2]);
  double lon = atan2(d[1], d[0]);
  double lat = asin(z);
  double c = (lon + OR_PI) * (1.0 / (2.0 * OR_PI)) * cols;
  double r = (OR_PI / 2.0 - lat) * (1.0 / OR_PI) * rows;
  int ic = (int)floor(c), ir = (int)floor(r);
  if (ic < 0) ic = 0;
  if (ic > cols - 1) ic = cols - 1;
  if (ir < 0) ir = 0;
  if (ir > rows - 1) ir = rows - 1;
  if (fc) *fc = c;
  if (fr) *fr = r;
  if (col) *col = ic;
  if (row) *row = ir;
}

/* ------------------------------------------------------------------------ */
/* Colour lookups                                                            */
/* ------------------------------------------------------------------------ */
static const bh_image* tex_at(const bh_image* t, int n, int slot) {
  if (slot < 0 || slot >= n || !t[slot].data) return NULL;
  return &t[slot];
}

static void texel_rgb(const bh_image* img, int col, int row, uint8_t rgb[3]) {
  const uint8_t* p = img->data + ((size_t)row * img->cols + col) * img->channels;
  rgb[0] = p[0];
  rgb[1] = p[1];
  rgb[2] = p[2];
}

/* Bilinear sample at continuous texel coordinates (a, b) = (col, row) with
 * texel centres at +0.5, clamped to the edge; result rounded to u8. */
static void bilinear_rgb(const bh_image* img, double a, double b, uint8_t rgb[3]) {
  double x = a - 0.5, y = b - 0.5;
  double fx = floor(x), fy = floor(y);
  double tx = x - fx, ty = y - fy;
  int x0 = (int)fx, y0 = (int)fy;
  int x1 = x0 + 1, y1 = y0 + 1;
  if (x0 < 0) x0 = 0;
  if (x0 > img->cols - 1) x0 = img->cols - 1;
  if (x1 < 0) x1 = 0;
  if (x1 > img->cols - 1) x1 = img->cols - 1;
  if (y0 < 0) y0 = 0;
  if (y0 > img->rows - 1) y0 = img->rows - 1;
  if (y1 < 0) y1 = 0;
  if (y1 > img->rows - 1) y1 = img->rows - 1;
  for (int ch = 0; ch < 3; ++ch) {
    const uint8_t* d = img->data;
    int C = img->channels;
    double v00 = d[((size_t)y0 * img->cols + x0) * C + ch];
    double v01 = d[((size_t)y0 * img->cols + x1) * C + ch];
    double v10 = d[((size_t)y1 * img->cols + x0) * C + ch];
    double v11 = d[((size_t)y1 * img->cols + x1) * C + ch];
    double top = v00 + (v01 - v00) * tx;
    double bot = v10 + (v11 - v10) * tx;
    double v = top + (bot - top) * ty;
    int iv = (int)floor(v + 0.5);
    rgb[ch] = (uint8_t)(iv < 0 ? 0 : (iv > 255 ? 255 : iv));
  }
}

/* ------------------------------------------------------------------------ */
/* FindCollision (object_manager.h:69-85): every object is tested and the    */
/* hit with the smallest |p1 - q|^2 wins; equal distances keep the object    */
/* tested first (hole, disc, spheres in order).                              */
/* ------------------------------------------------------------------------ */
typedef struct hit_rec {
  int cls;    /* BH_HIT_* */
  int index;  /* sphere index for BH_HIT_OBJECT */
  v3 q;
} hit_rec;

static int find_collision(const bh_scene* s, v3 p1, v3 p2, hit_rec* out) {
  int found = 0;
  double best = 0;
  v3 q;
  /* horizon: radius() = 2 * mass() */
  if (seg_sphere(p1, p2, ld(s->hole.position), 2 * s->hole.mass, &q)) {
    v3 d = sub(p1, q);
    double k = dot(d, d);
    best = k;
    found = 1;
    out->cls = BH_HIT_HORIZON;
    out->index = -1;
    out->q = q;
  }
  if (s->has_disc && annulus_collide(&s->disc, p1, p2, &q)) {
    v3 d = sub(p1, q);
    double k = dot(d, d);
    if (!found || k < best) {
      best = k;
      found = 1;
      out->cls = BH_HIT_DISC;
      out->index = -1;
      out->q = q;
    }
  }
  for (int i = 0; i < s->n_spheres; ++i) {
    const bh_sphere* sp = &s->spheres[i];
    if (seg_sphere(p1, p2, ld(sp->center), sp->radius, &q)) {
      v3 d = sub(p1, q);
      double k = dot(d, d);
      if (!found || k < best) {
        best = k;
        found = 1;
        out->cls = BH_HIT_OBJECT;
        out->index = i;
        out->q = q;
      }
    }
  }
  return found;
}

/* Colour of a hit: horizon black (blackhole_solution.h:61-63), disc texel
 * (vector_object.h:163-179), sphere texel (equirect of the normal). */
static void hit_color(const bh_scene* s, const bh_options* o, const bh_image* tex,
                      int ntex, const hit_rec* h, uint8_t rgb[3]) {
  rgb[0] = rgb[1] = rgb[2] = 0;
  if (h->cls == BH_HIT_DISC) {
    const bh_image* img = tex_at(tex, ntex, s->disc.texture);
    if (!img) return;
    double fw, fh;
    int pw, ph;
    or_rect_texel(s->disc.vertex, img->rows, img->cols, h->q.v, &fw, &fh, &pw, &ph);
    if (o->filter == BH_FILTER_BILINEAR)
      bilinear_rgb(img, fw, fh, rgb);
    else
      texel_rgb(img, pw, ph, rgb);
  } else if (h->cls == BH_HIT_OBJECT) {
    const bh_sphere* sp = &s->spheres[h->index];
    const bh_image* img = tex_at(tex, ntex, sp->texture);
    if (!img) return;
    v3 n = scl(sub(h->q, ld(sp->center)), 1.0 / sp->radius);
    int col, row;
    or_equirect_texel(n.v, img->rows, img->cols, NULL, NULL, &col, &row);
    texel_rgb(img, col, row, rgb);
  }
}

/* ------------------------------------------------------------------------ */
/* The march of one pixel (blackhole_solution_test.cc:161-298)               */
/* ------------------------------------------------------------------------ */
typedef struct march_state {
  double u, w_prev, phi;
  v3 P_prev, P_prev2;
  int steps;
} march_state;

/* One integration step: u already advanced by the caller. */
static int march_step(const bh_scene* s, const or_ray* ray, march_state* st, hit_rec* h) {
  double M = s->hole.mass;
  v3 B = ld(s->hole.position);
  double w = 1.0 / sqrt(or_G(st->u, ray->b, M));      /* InvSqrtG (:31-33)  */
  st->phi += (st->w_prev + w) * ray->du_h;              /* (:221)             */
  double r = 1. / st->u;                                /* (:222)             */
  v3 lv = mk(0, r * cos(st->phi), r * sin(st->phi));    /* (:224)             */
  v3 P = add(matvec(ray->R, lv), B);                    /* (:225)             */
  st->w_prev = w;
  st->steps++;
  if (find_collision(s, st->P_prev, P, h)) return 1;
  st->P_prev2 = st->P_prev;
  st->P_prev = P;
  return 0;
}

static void write_px(const bh_outputs* out, size_t idx, const uint8_t rgb[3], int cls,
                     double phi, const v3* q, int steps) {
  if (out->rgb) {
    out->rgb[3 * idx + 0] = rgb[0];
    out->rgb[3 * idx + 1] = rgb[1];
    out->rgb[3 * idx + 2] = rgb[2];
  }
  if (out->hit) out->hit[idx] = (uint8_t)cls;
  if (out->phi_end) out->phi_end[idx] = phi;
  if (out->hit_point) {
    for (int k = 0; k < 3; ++k) out->hit_point[3 * idx + k] = q ? q->v[k] : NAN;
  }
  if (out->steps) out->steps[idx] = steps;
}

/* The march of pixel (x, y).  chord, if not NULL, receives the last two
 * points of the march (P_prev2 then P_prev) of a ray that hits nothing: the
 * chord whose direction picks the background texel. */
static int trace_pixel(const bh_scene* s, const bh_options* o, const bh_image* tex, int ntex,
                       int x, int y, const bh_outputs* out, double chord[6]) {
  const int W = s->camera.width;
  size_t idx = (size_t)y * W + x;
  uint8_t black[3] = {0, 0, 0};
  or_ray ray;
  or_ray_setup(s, o, x, y, &ray);
  if (ray.degenerate) {
    write_px(out, idx, black, BH_HIT_HORIZON, NAN, NULL, 0);
    return 0;
  }
  int nstep = o->nstep > 0 ? o->nstep : 20;
  march_state st;
  st.u = ray.u0;
  st.w_prev = 0;
  st.phi = ray.phi0;
  st.P_prev = ld(s->camera.position);
  st.P_prev2 = st.P_prev;
  st.steps = 0;
  hit_rec h;
  uint8_t rgb[3];

  /* inbound: nstep - 1 steps of +du (:217-239) */
  for (int i = 0; i < nstep - 1; ++i) {
    st.u += ray.du;
    if (march_step(s, &ray, &st, &h)) goto hit;
  }
  /* final approach: +0.9 du, still weighted du/2 (:240-262) */
  st.u += ray.du * 0.9;
  if (march_step(s, &ray, &st, &h)) goto hit;

  /* captured: straight chord to the centre (:264-272) */
  if (ray.captured) {
    if (find_collision(s, st.P_prev, ld(s->hole.position), &h)) goto hit;
    write_px(out, idx, black, BH_HIT_NONE, st.phi, NULL, st.steps);
    return 0;
  }
  /* outbound: nstep - 1 steps of -du (:274-296) */
  for (int i = 0; i < nstep - 1; ++i) {
    st.u -= ray.du;
    if (march_step(s, &ray, &st, &h)) goto hit;
  }
  if (o->escape) {
    /* extension: keep stepping outward until the next point would lie at
     * r >= r_escape (or u <= 0), or the step cap is reached */
    while (st.steps < o->step_cap) {
      double un = st.u - ray.du;
      if (un * o->r_escape <= 1.0) break;
      st.u = un;
      if (march_step(s, &ray, &st, &h)) goto hit;
    }
    if (chord)
      for (int k = 0; k < 3; ++k) {
        chord[k] = st.P_prev2.v[k];
        chord[3 + k] = st.P_prev.v[k];
      }
    const bh_image* bg = tex_at(tex, ntex, s->background);
    uint8_t c[3] = {0, 0, 0};
    if (bg) {
      v3 d = normalize(sub(st.P_prev, st.P_prev2));
      int col, row;
      or_equirect_texel(d.v, bg->rows, bg->cols, NULL, NULL, &col, &row);
      texel_rgb(bg, col, row, c);
    }
    write_px(out, idx, c, BH_HIT_NONE, st.phi, NULL, st.steps);
    return 0;
  }
  write_px(out, idx, black, BH_HIT_NONE, st.phi, NULL, st.steps);
  return 0;

hit:
  hit_color(s, o, tex, ntex, &h, rgb);
  write_px(out, idx, rgb, h.cls, st.phi, &h.q, st.steps);
  return 0;
}

int or_trace_pixel(const bh_scene* s, const bh_options* o, const bh_image* tex, int ntex,
                   int x, int y, const bh_outputs* out) {
  return trace_pixel(s, o, tex, ntex, x, y, out, NULL);
}

int or_trace_escape(const bh_scene* s, const bh_options* o, const bh_image* tex, int ntex,
                    int x, int y, const bh_outputs* out, double chord[6]) {
  for (int k = 0; k < 6; ++k) chord[k] = NAN;
  return trace_pixel(s, o, tex, ntex, x, y, out, chord);
}

/* ------------------------------------------------------------------------ */
/* Frames                                                                    */
/* ------------------------------------------------------------------------ */
typedef struct job {
  const bh_scene* s;
  const bh_options* o;
  const bh_image* tex;
  int ntex, y0, y1, tid, nthreads;
  const bh_outputs* out;
} job;

static void* run_rows(void* arg) {
  job* j = (job*)arg;
  int W = j->s->camera.width;
  for (int y = j->y0 + j->tid; y < j->y1; y += j->nthreads)
    for (int x = 0; x < W; ++x) or_trace_pixel(j->s, j->o, j->tex, j->ntex, x, y, j->out);
  return NULL;
}

int or_render_rows(const bh_scene* s, const bh_options* o, const bh_image* tex, int ntex,
                   int threads, int y0, int y1, const bh_outputs* out) {
  if (!s || !o || !out) return BH_ERR_INVALID_ARG;
  if (threads < 1) threads = 1;
  if (threads > 256) threads = 256;
  pthread_t th[256];
  job jobs[256];
  for (int t = 0; t < threads; ++t) {
    jobs[t] = (job){s, o, tex, ntex, y0, y1, t, threads, out};
    if (threads == 1) {
      run_rows(&jobs[t]);
    } else if (pthread_create(&th[t], NULL, run_rows, &jobs[t]) != 0) {
      return BH_ERR_INVALID_ARG;
    }
  }
  if (threads > 1)
    for (int t = 0; t < threads; ++t) pthread_join(th[t], NULL);
  return 0;
}

int or_render(const bh_scene* s, const bh_options* o, const bh_image* tex, int ntex,
              int threads, const bh_outputs* out) {
  return or_render_rows(s, o, tex, ntex, threads, 0, s->camera.height, out);
}

/* Inbound sweep for an arbitrary (b, M, u0): phi after the final approach
 * step minus phi0.  Same loop as or_trace_pixel without geometry. */
double or_inbound_sweep(double b, double M, double u0, int nstep, int root_mode) {
  double bc = or_b_c(M);
  double u_end = b < bc ? 1.0 / (3 * M) : or_solve_g(b, M, root_mode);
  double du = (u_end - u0) * (1.0 / nstep);
  double du_h = du / 2.;
  double u = u0, wp = 0, phi = 0;
  for (int i = 0; i < nstep - 1; ++i) {
    u += du;
    double w = 1.0 / sqrt(or_G(u, b, M));
    phi += (wp + w) * du_h;
    wp = w;
  }
  u += du * 0.9;
  double w = 1.0 / sqrt(or_G(u, b, M));
  phi += (wp + w) * du_h;
  return phi;
}

/* ------------------------------------------------------------------------ */
/* Differentiable mode (extension)                                           */
/* ------------------------------------------------------------------------ */
void or_diff_scene(const bh_diff_params* p, bh_scene* s, bh_options* o) {
  memset(s, 0, sizeof *s);
  memset(o, 0, sizeof *o);
  double si = sin(p->incl), ci = cos(p->incl);
  v3 C = mk(p->dist * si * cos(p->azim), p->dist * si * sin(p->azim), p->dist * ci);
  /* look-at the origin with +z up: x toward the hole, z = -(up orthogonal), y = z x x */
  v3 x = normalize(scl(C, -1.0));
  v3 up = mk(0, 0, 1);
  double k = dot(up, x);
  v3 z = normalize(scl(sub(up, scl(x, k)), -1.0));
  v3 y = normalize(cross(z, x));
  memcpy(s->camera.position, C.v, sizeof C.v);
  memcpy(s->camera.axis_x, x.v, sizeof x.v);
  memcpy(s->camera.axis_y, y.v, sizeof y.v);
  memcpy(s->camera.axis_z, z.v, sizeof z.v);
  s->camera.width = p->width;
  s->camera.height = p->height;
  s->camera.fov = p->fov;
  s->hole.mass = p->r_s / 2.0;
  s->has_disc = 1;
  double R = p->r_outer;
  double P[5][3] = {{0, 0, 0}, {R, R, 0}, {-R, R, 0}, {-R, -R, 0}, {R, -R, 0}};
  memcpy(s->disc.vertex, P, sizeof P);
  v3 n = normalize(cross(sub(ld(P[3]), ld(P[1])), sub(ld(P[2]), ld(P[1]))));
  memcpy(s->disc.normal, n.v, sizeof n.v);
  s->disc.r_outer = p->r_outer;
  s->disc.r_inner = p->r_inner;
  s->disc.texture = p->texture;
  s->background = -1;
  o->precision = BH_PREC_FP64;
  o->nstep = 20;
  o->root = BH_ROOT_EXACT;
  o->filter = BH_FILTER_BILINEAR;
  o->edge_width = p->edge_width;
}

/* smoothstep of clamp(x + 1/2, 0, 1) */
static double soft_step(double x) {
  double t = x + 0.5;
  t = t < 0 ? 0 : (t > 1 ? 1 : t);
  return t * t * (3.0 - 2.0 * t);
}

static double soft_alpha(const bh_annulus* a, double rho, double w) {
  if (w <= 0) return (rho > a->r_outer || rho < a->r_inner) ? 0.0 : 1.0;
  return soft_step((rho - a->r_inner) / w) * soft_step((a->r_outer - rho) / w);
}

/* bilinear sample (texel centres at +0.5, clamped), channels as doubles;
 * *cell receives the clamped top-left texel index (col + cols * row) */
static void bilinear_d(const bh_image* img, double a, double b, double out[3], int* cell) {
  double x = a - 0.5, y = b - 0.5;
  double fx = floor(x), fy = floor(y);
  double tx = x - fx, ty = y - fy;
  int x0 = (int)fx, y0 = (int)fy, x1 = x0 + 1, y1 = y0 + 1;
  int W = img->cols, H = img->rows, C = img->channels;
  x0 = x0 < 0 ? 0 : (x0 > W - 1 ? W - 1 : x0);
  x1 = x1 < 0 ? 0 : (x1 > W - 1 ? W - 1 : x1);
  y0 = y0 < 0 ? 0 : (y0 > H - 1 ? H - 1 : y0);
  y1 = y1 < 0 ? 0 : (y1 > H - 1 ? H - 1 : y1);
  for (int ch = 0; ch < 3; ++ch) {
    double v00 = img->data[((size_t)y0 * W + x0) * C + ch];
    double v01 = img->data[((size_t)y0 * W + x1) * C + ch];
    double v10 = img->data[((size_t)y1 * W + x0) * C + ch];
    double v11 = img->data[((size_t)y1 * W + x1) * C + ch];
    double top = v00 + (v01 - v00) * tx;
    double bot = v10 + (v11 - v10) * tx;
    out[ch] = top + (bot - top) * ty;
  }
  *cell = x0 + W * y0;
}

/* FNV-1a style fold of one value into a structure signature */
static int32_t sig_mix(int32_t h, int32_t v) {
  return (int32_t)(((uint32_t)h ^ (uint32_t)v) * 16777619u);
}

/* continuous texel coordinates of a disc point: r cos(theta) = s1.v/|s1|,
 * r sin(theta) = |s1 x v|/|s1| (the smooth form of Rectangle::color) */
static void disc_coords(const bh_annulus* a, const bh_image* img, v3 q, double* U, double* V) {
  v3 v1 = ld(a->vertex[1]);
  v3 s1 = sub(ld(a->vertex[2]), v1);
  v3 s2 = sub(ld(a->vertex[4]), v1);
  v3 v = sub(q, v1);
  double l1 = sqrt(dot(s1, s1)), l2 = sqrt(dot(s2, s2));
  v3 c = cross(s1, v);
  *U = dot(s1, v) / (l1 * l1) * img->cols;
  *V = sqrt(dot(c, c)) / (l1 * l2) * img->rows;
}

typedef struct diff_acc {
  double T;        /* transmittance */
  double rgb[3];
  int n, first_cls;
  int32_t sig;
  int done;
} diff_acc;

static void diff_disc(const bh_scene* s, const bh_image* img, double w, v3 q, int step,
                      diff_acc* acc) {
  double rho = sqrt(dot(q, q));
  double al = soft_alpha(&s->disc, rho, w);
  if (al <= 0) return;
  double c[3] = {0, 0, 0};
  int cell = -1;
  if (img) {
    double U, V;
    disc_coords(&s->disc, img, q, &U, &V);
    bilinear_d(img, U, V, c, &cell);
  }
  for (int k = 0; k < 3; ++k) acc->rgb[k] += acc->T * al * c[k];
  acc->T *= 1.0 - al;
  if (!acc->first_cls) acc->first_cls = BH_HIT_DISC;
  acc->sig = sig_mix(sig_mix(acc->sig, step + 1), cell);
  acc->n++;
  if (acc->T <= 0) acc->done = 1;
}

/* One chord of the differentiable march: disc crossing (soft window) and
 * horizon, nearer first. */
static void diff_chord(const bh_scene* s, const bh_image* img, double w, v3 p1, v3 p2, int step,
                       diff_acc* acc) {
  v3 B = ld(s->hole.position);
  v3 qh;
  int hh = seg_sphere(p1, p2, B, 2 * s->hole.mass, &qh);
  double dh = 0;
  if (hh) {
    v3 d = sub(p1, qh);
    dh = dot(d, d);
  }
  /* disc crossing without the radius window */
  v3 c0 = ld(s->disc.vertex[0]);
  v3 n = ld(s->disc.normal);
  v3 a1 = sub(p1, c0), a2 = sub(p2, c0);
  int hd = 0;
  v3 qd;
  double dd = 0;
  if (dot(n, a1) * dot(n, a2) < 0) {
    double t1 = dot(n, a1), t2 = dot(n, a2);
    v3 c = scl(add(scl(a1, fabs(t2)), scl(a2, fabs(t1))), 1. / (fabs(t1) + fabs(t2)));
    if (soft_alpha(&s->disc, sqrt(dot(c, c)), w) > 0) {
      hd = 1;
      qd = add(c, c0);
      v3 d = sub(p1, qd);
      dd = dot(d, d);
    }
  }
  if (hd && (!hh || dd < dh)) {
    diff_disc(s, img, w, qd, step, acc);
    if (acc->done) return;
  }
  if (hh) {
    if (!acc->first_cls) acc->first_cls = BH_HIT_HORIZON;
    acc->T = 0;
    acc->done = 1;
    return;
  }
}

static void diff_pixel(const bh_scene* s, const bh_options* o, const bh_image* img, double w,
                       int x, int y, double* rgb, uint8_t* hit, int32_t* sig) {
  or_ray ray;
  or_ray_setup(s, o, x, y, &ray);
  diff_acc acc = {1.0, {0, 0, 0}, 0, 0, (int32_t)2166136261u, 0};
  int steps = 0;
  if (ray.degenerate) {
    acc.first_cls = BH_HIT_HORIZON;
    goto out;
  }
  {
    int nstep = o->nstep;
    double M = s->hole.mass;
    v3 B = ld(s->hole.position);
    double u = ray.u0, wp = 0, phi = ray.phi0;
    v3 Pp = ld(s->camera.position);
#define DSTEP()                                                   \
    do {                                                          \
      double wn = 1.0 / sqrt(or_G(u, ray.b, M));                  \
      phi += (wp + wn) * ray.du_h;                                \
      double r = 1. / u;                                          \
      v3 P = add(matvec(ray.R, mk(0, r * cos(phi), r * sin(phi))), B); \
      wp = wn;                                                    \
      ++steps;                                                    \
      diff_chord(s, img, w, Pp, P, steps, &acc);                  \
      Pp = P;                                                     \
    } while (0)
    for (int i = 0; i < nstep - 1 && !acc.done; ++i) {
      u += ray.du;
      DSTEP();
    }
    if (!acc.done) {
      u += ray.du * 0.9;
      DSTEP();
    }
    if (!acc.done && ray.captured) {
      /* chord to the centre: the disc sees n.(B - B) = 0, the horizon is entered */
      if (!acc.first_cls) acc.first_cls = BH_HIT_HORIZON;
      acc.T = 0;
      acc.done = 1;
    }
    for (int i = 0; i < nstep - 1 && !acc.done; ++i) {
      u -= ray.du;
      DSTEP();
    }
#undef DSTEP
  }
out:
  if (rgb)
    for (int k = 0; k < 3; ++k) rgb[k] = acc.rgb[k];
  if (hit) *hit = (uint8_t)acc.first_cls;
  if (sig) *sig = sig_mix(sig_mix(acc.sig, steps), acc.first_cls);
}

typedef struct diff_job {
  const bh_scene* s;
  const bh_options* o;
  const bh_image* img;
  double w;
  int tid, nthreads;
  double* rgb;
  uint8_t* hit;
  int32_t* sig;
} diff_job;

static void* diff_rows(void* arg) {
  diff_job* j = (diff_job*)arg;
  int W = j->s->camera.width, H = j->s->camera.height;
  for (int y = j->tid; y < H; y += j->nthreads)
    for (int x = 0; x < W; ++x) {
      size_t idx = (size_t)y * W + x;
      diff_pixel(j->s, j->o, j->img, j->w, x, y, j->rgb ? j->rgb + 3 * idx : NULL,
                 j->hit ? j->hit + idx : NULL, j->sig ? j->sig + idx : NULL);
    }
  return NULL;
}

int or_render_diff(const bh_diff_params* p, const bh_image* tex, int ntex, int threads,
                   double* rgb, uint8_t* hit, int32_t* sig) {
  if (!p || p->width <= 0 || p->height <= 0) return BH_ERR_INVALID_ARG;
  bh_scene s;
  bh_options o;
  or_diff_scene(p, &s, &o);
  const bh_image* img = tex_at(tex, ntex, p->texture);
  if (threads < 1) threads = 1;
  if (threads > 256) threads = 256;
  pthread_t th[256];
  diff_job jobs[256];
  for (int t = 0; t < threads; ++t) {
    jobs[t] = (diff_job){&s, &o, img, p->edge_width, t, threads, rgb, hit, sig};
    if (threads == 1)
      diff_rows(&jobs[t]);
    else if (pthread_create(&th[t], NULL, diff_rows, &jobs[t]) != 0)
      return BH_ERR_INVALID_ARG;
  }
  if (threads > 1)
    for (int t = 0; t < threads; ++t) pthread_join(th[t], NULL);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Flat march: RayTracer::Prograde (ray_tracer.h:17-96)                      */
/* ------------------------------------------------------------------------ */
/* Rectangle::Collide (vector_object.h:107-127) */
static int rect_collide(const bh_rect* r, v3 aa, v3 bb, v3* inter) {
  v3 p0 = ld(r->vertex[1]);
  v3 p1 = sub(ld(r->vertex[2]), p0);
  v3 p3 = sub(ld(r->vertex[4]), p0);
  v3 n = normalize(cross(p1, p3));
  v3 Q1 = sub(aa, p0);
  v3 Q2 = sub(bb, p0);
  double q1 = dot(n, Q1);
  double q2 = dot(n, Q2);
  if (q1 * q2 >= 0) return 0;
  v3 Q = scl(add(scl(Q1, fabs(q2)), scl(Q2, fabs(q1))), 1. / (fabs(q1) + fabs(q2)));
  if (dot(p1, Q) > 0 && dot(p1, p1) > dot(p1, Q) && dot(p3, Q) > 0 && dot(p3, p3) > dot(p3, Q)) {
    *inter = add(Q, p0);
    return 1;
  }
  return 0;
}

/* InfinitePlane::Collide (vector_object.h:210-225) */
static int plane_collide(const bh_plane* pl, v3 p1, v3 p2, v3* inter) {
  const double* z = pl->axis_z;
  const double* o = pl->position;
  double t1 = z[0] * (p1.v[0] - o[0]) + z[1] * (p1.v[1] - o[1]) + z[2] * (p1.v[2] - o[2]);
  double t2 = z[0] * (p2.v[0] - o[0]) + z[1] * (p2.v[1] - o[1]) + z[2] * (p2.v[2] - o[2]);
  if (t1 * t2 > 0) return 0;
  double d1 = fabs(t1), d2 = fabs(t2);
  *inter = scl(add(scl(p1, d2), scl(p2, d1)), 1. / (d1 + d2));
  return 1;
}

/* ChessPattern2D (pattern.h:14-50) */
/* ChessPattern2D::mod (pattern.h:44-47).  Beyond the int range, where the
 * reference's (int) cast is undefined, the same truncated quotient is taken
 * in double arithmetic. */
static double chess_mod(double a, double ps) {
  double b = fabs(a);
  if (!(b < 2147483648.0)) return b - floor(floor(b) / (ps * 2)) * (ps * 2);
  return b - ((int)(b) / (int)(ps * 2)) * (ps * 2);
}

static void chess(double ps, double x, double y, uint8_t rgb[3]) {
  double x2 = chess_mod(x, ps), y2 = chess_mod(y, ps);
  int in = (x2 <= ps && y2 <= ps) || (ps <= x2 && ps <= y2);
  int white = (x * y > 0) ? in : !in;
  rgb[0] = rgb[1] = rgb[2] = white ? 255 : 0;
}

/* InfinitePlane::color (vector_object.h:227-232) */
static void plane_color(const bh_plane* pl, v3 q, uint8_t rgb[3]) {
  const double* vx = pl->axis_x;
  const double* vy = pl->axis_y;
  double x = q.v[0], y = q.v[1];
  double b = (vx[0] * y - vx[1] * x) / (vx[0] * vy[1] - vx[1] * vy[0]);
  double a = (x - b * vy[0]) / vx[0];
  if (pl->pattern_size > 0)
    chess(pl->pattern_size, a, b, rgb);
  else
    rgb[0] = rgb[1] = rgb[2] = 0;
}

int or_rect_collide(const bh_rect* r, const double a[3], const double b[3], double out[3]) {
  v3 q;
  int h = rect_collide(r, ld(a), ld(b), &q);
  if (h) memcpy(out, q.v, sizeof q.v);
  return h;
}

int or_plane_collide(const bh_plane* pl, const double a[3], const double b[3], double out[3]) {
  v3 q;
  int h = plane_collide(pl, ld(a), ld(b), &q);
  if (h) memcpy(out, q.v, sizeof q.v);
  return h;
}

void or_plane_color(const bh_plane* pl, const double q[3], uint8_t rgb[3]) {
  plane_color(pl, ld(q), rgb);
}

int or_render_flat(const bh_flat_scene* s, const bh_image* tex, int ntex,
                   const bh_outputs* out) {
  const bh_camera* cam = &s->camera;
  v3 fv = focus_vector(cam);
  for (int x = 0; x < cam->width; ++x) {
    for (int y = 0; y < cam->height; ++y) {
      size_t idx = (size_t)y * cam->width + x;
      v3 old = ld(cam->position);
      v3 present = pixel_vector(cam, fv, x, y);
      /* BasicLinearRayRecurrence (ray_tracer.h:21-27) */
      v3 vec = sub(present, old);
      double d2 = dot(vec, vec);
      if (d2 < 100) vec = scl(vec, 100.0 / d2);
      uint8_t rgb[3] = {0, 0, 0};
      int cls = BH_HIT_NONE, st = 0;
      v3 q = mk(NAN, NAN, NAN);
      for (int k = 0; k < s->steps; ++k) {
        st++;
        int found = 0, kind = -1, which = -1;
        double best = 0;
        v3 c;
        for (int i = 0; i < s->n_rects; ++i) {
          if (rect_collide(&s->rects[i], old, present, &c)) {
            v3 d = sub(old, c);
            double kk = dot(d, d);
            if (!found || kk < best) { best = kk; found = 1; kind = 0; which = i; q = c; }
          }
        }
        for (int i = 0; i < s->n_planes; ++i) {
          if (plane_collide(&s->planes[i], old, present, &c)) {
            v3 d = sub(old, c);
            double kk = dot(d, d);
            if (!found || kk < best) { best = kk; found = 1; kind = 1; which = i; q = c; }
          }
        }
        if (found) {
          cls = BH_HIT_OBJECT;
          if (kind == 0) {
            const bh_image* img = tex_at(tex, ntex, s->rects[which].texture);
            if (img) {
              int pw, ph;
              or_rect_texel(s->rects[which].vertex, img->rows, img->cols, q.v, NULL, NULL,
                            &pw, &ph);
              texel_rgb(img, pw, ph, rgb);
            }
          } else {
            plane_color(&s->planes[which], q, rgb);
          }
          break;
        }
        v3 np = add(present, vec);
        old = present;
        present = np;
      }
      write_px(out, idx, rgb, cls, NAN, cls ? &q : NULL, st);
    }
  }
  return 0;
}
