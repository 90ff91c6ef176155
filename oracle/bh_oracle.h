/*
 * bh_oracle.h — FP64 CPU oracle for the photon-geodesic march.
 *
 * TEST INFRASTRUCTURE ONLY.  Results are checked against it: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and
 * never as the thing measured or shipped.  The product path is the HIP
 * library under blackhole_8_amd/, which never links or calls this code.
 *
 * It restates, in plain C with IEEE double arithmetic in the reference's
 * operation order (compiled with -ffp-contract=off), the reference march of
 * include/blackhole/blackhole_solution_test.cc:161-298 and the objects it
 * calls.  Scene, option and output types are the C ABI's (include/bh_api.h)
 * so the same inputs drive the oracle and the GPU.
 */
#ifndef BH_ORACLE_H_
#define BH_ORACLE_H_

#include "../include/bh_api.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Per-ray set-up (blackhole_solution_test.cc:167-211). */
typedef struct or_ray {
  double pv[3], F[3];          /* pixel point and camera, hole-relative      */
  double R[9];                 /* [zv yv xv] as columns, row-major            */
  double Rinv[9];              /* Matx33::inv() (adjugate / determinant)      */
  double Fp[3];                /* Rinv * F                                    */
  double b, r0, phi0, u0;      /* impact distance, |F'|, atan(b/Y0), 1/r0     */
  double u_end;                /* SolveG(b) or 1/(3M) when captured           */
  double du, du_h;
  int captured;                /* b < b_c                                     */
  int degenerate;              /* orbit-plane frame singular (ray through B)  */
} or_ray;

double or_G(double u, double b, double M);
double or_b_c(double M);
double or_solve_g(double b, double M, int root_mode);
double or_focal_length(int width, double fov);
void or_pixel_vector(const bh_camera* cam, int x, int y, double out[3]);
void or_inv3(const double a[9], double out[9], int* ok);
int or_ray_setup(const bh_scene* s, const bh_options* o, int x, int y, or_ray* ray);

/* Object tests (return 1 on hit and write the intersection). */
int or_seg_sphere(const double q1[3], const double q2[3], const double c[3], double R,
                  double inter[3]);
int or_annulus_collide(const bh_annulus* a, const double q1[3], const double q2[3],
                       double inter[3]);
/* Rectangle::color texel indices (vector_object.h:163-179), clamped. */
void or_rect_texel(const double vtx[5][3], int rows, int cols, const double p[3],
                   double* fw, double* fh, int* pw, int* ph);
/* Equirectangular lookup of a unit direction: continuous and integer coords. */
void or_equirect_texel(const double d[3], int rows, int cols, double* fc, double* fr,
                       int* col, int* row);

/* One pixel.  textures[slot] are host images.  Writes one pixel of out. */
int or_trace_pixel(const bh_scene* s, const bh_options* o, const bh_image* textures,
                   int n_textures, int x, int y, const bh_outputs* out);
/* or_trace_pixel that also returns, for a ray that escapes with the escape
 * extension on, the last two points of its march (chord[0..2] = the point
 * before the last, chord[3..5] = the last); NaN otherwise. */
int or_trace_escape(const bh_scene* s, const bh_options* o, const bh_image* textures,
                    int n_textures, int x, int y, const bh_outputs* out, double chord[6]);
/* Whole frame on `threads` host threads (rows interleaved). */
int or_render(const bh_scene* s, const bh_options* o, const bh_image* textures,
              int n_textures, int threads, const bh_outputs* out);
/* Rows [y0, y1) only, for bounded timing samples. */
int or_render_rows(const bh_scene* s, const bh_options* o, const bh_image* textures,
                   int n_textures, int threads, int y0, int y1, const bh_outputs* out);

/* Inbound trapezoid phase only: phi after the final approach step minus
 * phi0, for convergence checks against the exact integral. */
double or_inbound_sweep(double b, double M, double u0, int nstep, int root_mode);

/* Differentiable mode (no reference counterpart): the march of
 * bh_diff_params with the exact turning point, bilinear texels and soft disc
 * edges composited front to back.  rgb: H*W*3 doubles in [0, 255]; hit:
 * class of the first contribution; sig: per-pixel hash of the discrete
 * structure (step and bilinear texel cell of every contribution, the
 * terminating step and the class), for masking finite differences.  Any output may be NULL. */
int or_render_diff(const bh_diff_params* p, const bh_image* textures, int n_textures,
                   int threads, double* rgb, uint8_t* hit, int32_t* sig);
/* The scene and options the differentiable mode renders (reference march). */
void or_diff_scene(const bh_diff_params* p, bh_scene* s, bh_options* o);

/* Flat march (ray_tracer.h:52-96) and its object tests: Rectangle::Collide
 * (vector_object.h:107-127), InfinitePlane::Collide / color with
 * ChessPattern2D (vector_object.h:210-232, pattern.h:14-50). */
int or_rect_collide(const bh_rect* r, const double a[3], const double b[3], double out[3]);
int or_plane_collide(const bh_plane* pl, const double a[3], const double b[3], double out[3]);
void or_plane_color(const bh_plane* pl, const double q[3], uint8_t rgb[3]);
int or_render_flat(const bh_flat_scene* s, const bh_image* textures, int n_textures,
                   const bh_outputs* out);

#ifdef __cplusplus
}
#endif
#endif
