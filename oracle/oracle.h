/*
 * oracle.h — TEST INFRASTRUCTURE ONLY.
 *
 * The oracle of the path-tracing hot path: a float32 CPU restatement of the
 * reference CPURenderer integrators (M/renderer/cpu/AbstractCPURenderer.java
 * :170-258 over M/scene/Scene.java:272-326, 451-718, 797-900, 1066-1090,
 * 1530-1633) and an FP64 closest-hit intersector.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it; the
 * product never links or calls it.
 *
 * How each part is pinned (tests/test_oracle_*.py):
 *   triangle test   reference golden vectors are absent for Triangle3F; pinned
 *                   by the FP64 brute force, closed-form hits, watertightness
 *                   on closed meshes and shared edges
 *   sphere / plane  T/geometry/shape/Sphere3DUnitTests.java:149-197 and
 *                   T/geometry/shape/Plane3FUnitTests.java:190-215 vectors
 *   camera          closed form (centre pixel -> W, corners)
 *   Fresnel         closed forms ((eta-1)/(eta+1))^2, conductor at normal
 *                   incidence, total internal reflection
 *   BxDFs           white furnace (energy <= 1), reciprocity, pdf integrates
 *                   to one, sampled-direction histograms against the pdf
 *   integrators     analytic furnace scene (Le/(1-a)), AO of an open plane
 *                   and of a box interior in closed form
 *   Perez sky       zenith luminance and chromaticity in closed form,
 *                   symmetry, horizon, pdf normalisation
 *   Perez sun       attenuated spectrum against Preetham et al.'s
 *                   transmittance formulas in float64, colour against an
 *                   independent CIE integration and sRGB matrix
 *   transforms      identity matrices change no bit; a rigid motion of
 *                   every primitive and the camera leaves each path the same
 *                   (frames built in object space), which moving the
 *                   vertices themselves does not; the shading normal of a
 *                   non-uniformly scaled mesh equals the normal matrix's
 *                   image in closed form (tests/test_transforms.py).  The hit
 *                   is found on the primitive's world-space image, with the
 *                   0.001 minimum distance in world space (the reference
 *                   applies it to the object-space ray)
 *   image light     a constant image: constant radiance and a pdf that
 *                   integrates to one; texel lookup at directions whose image
 *                   coordinates are known in closed form (wrap-around at the
 *                   seam, the object frame, rotation and scale); samples carry
 *                   the evaluated pdf and follow it; a Lambertian plane under
 *                   a constant sky renders to albedo x radiance
 *                   (tests/test_image_light.py)
 *   CPU film        filter table against the Gaussian in closed form; a
 *                   splat covers exactly the pixels within the radius, with
 *                   weights symmetric and largest at the nearest centre; the
 *                   sRGB matrices against IEC 61966-2-1's; the gamma
 *                   constants against the standard curve (12.92, 1.055,
 *                   0.055); a constant-radiance furnace film renders to the
 *                   encoded constant (tests/test_film.py)
 *   image stage     textbook sRGB bytes (linear 0.5 -> 188, 0.18 -> 118),
 *                   tone curves at closed-form points (Reinhard 1 -> 1/2,
 *                   white point 4 -> 1, 1 - e^-ln2 = 1/2, ACES fit at 1),
 *                   composition order (tone map before the sRGB encode)
 *
 * M/ = src/main/java/org/dayflower/ and T/ = src/test/java/org/dayflower/ of
 * the reference tree.
 */
#ifndef DAYFLOWER_ORACLE_H
#define DAYFLOWER_ORACLE_H

#include <stdint.h>
#include "../include/dayflower_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_scene orc_scene;

typedef struct orc_params {
    int32_t width, height, spp, spp_offset;
    int32_t algorithm, max_bounce, min_rr, samples;
    int32_t light_pdf_mode, threads;
    int32_t x0, y0, x1, y1;      /* pixel window; x1/y1 <= 0 means full    */
    float max_distance;
    int32_t preview;             /* setPreviewMode: a primary miss is white */
    uint64_t seed;
} orc_params;

orc_scene* orc_scene_create(const dfr_scene_desc* desc);
void orc_scene_destroy(orc_scene* scene);
int32_t orc_scene_set_camera(orc_scene* scene, const float* camera19);
/* Per-primitive object-to-world matrices (16 floats each, row major); the
 * records the scene was created with are the object-space forms.  Returns
 * -1 on a count mismatch, a non-affine or a singular matrix. */
int32_t orc_scene_set_transforms(orc_scene* scene, const float* m, int32_t n);

/* Per-pixel sum and sum of squares of the accepted samples and their count
 * (unfiltered film, one bin per pixel).  Arrays are width*height*3 doubles
 * and width*height ints; any may be NULL. */
int32_t orc_render(const orc_scene* scene, const orc_params* params, double* sum, double* sumsq,
                   int32_t* count);

/* CPURenderer's film (PixelImageF with GaussianFilter2F(2, 2, 2), sRGB):
 * orc_render_film accumulates `params->spp` passes into film (W*H*4 floats:
 * X, Y, Z, filter weight sum; the caller zeroes it); orc_film_render turns
 * it into the displayed gamma-encoded RGB (filmRender); orc_film_splat is
 * filmAddColorXYZ of one sample; orc_film_constants returns the 16 x 16
 * filter table, the XYZ->RGB and RGB->XYZ arrays (12 floats each, as the
 * reference stores them) and the gamma constants (break point, slope, slope
 * match, segment offset).  Any output may be NULL. */
int32_t orc_render_film(const orc_scene* scene, const orc_params* params, float* film);
/* An ImageLight alone (tests): radiance and pdf along n world directions;
 * samples for n_samples (u, v) pairs (direction, pdf, radiance; pdf 0 when
 * there is no sample). */
int32_t orc_image_light(int32_t w, int32_t h, const float* rgb, float angle_degrees, float su, float sv, int64_t n,
                        const float* dirs, float* radiance, float* pdf, int64_t n_samples, const float* uv,
                        float* sample_dirs, float* sample_pdf, float* sample_radiance);
int32_t orc_film_render(int64_t n, const float* film, float* rgb);
int32_t orc_film_splat(int32_t width, int32_t height, float x, float y, const float* xyz, float* film);
int32_t orc_film_constants(float* table256, float* xyz_to_rgb12, float* rgb_to_xyz12, float* gamma4);

/* Closest/any hit for rays of 8 floats (origin, tmin, direction, tmax). */
int32_t orc_trace_f32(const orc_scene* scene, const float* rays, int64_t n, int32_t any_hit,
                      int32_t threads, int32_t* prim, int64_t* tri, float* t);
int32_t orc_trace_f64(const orc_scene* scene, const float* rays, int64_t n, int32_t any_hit,
                      int32_t threads, int32_t* prim, int64_t* tri, double* t, int64_t* nodes,
                      int64_t* tris);
/* FP64 walk of the product's four-wide BVH (dfr_read_bvh layout) over the
 * 12-float ray records of dfr_read_ray_sample: node visits, triangle tests,
 * hit triangle id (-1: miss) and distance per ray (the last two may be NULL). */
int32_t orc_bvh4_walk(const float* nodes, int64_t n_nodes, const float* tris, int64_t n_tris, const float* rays,
                      int64_t n, int32_t threads, int64_t* node_visits, int64_t* tri_tests, int64_t* hit_tri,
                      double* hit_t);
int32_t orc_primary_rays(const orc_scene* scene, int32_t width, int32_t height, uint64_t seed,
                         int32_t sample, float* rays);
int32_t orc_camera_ray(const float* camera19, float ix, float iy, float px, float py, float* out6);
/* the same with the thin lens's disk sample (lu, lv) */
int32_t orc_lens_ray(const float* camera19, float ix, float iy, float px, float py, float lu, float lv, float* out6);
/* Tests only: 1 makes the float triangle test compare t * det with tmin
 * literally, as the reference does (Triangle3F.java:570-574). */
void orc_set_literal_tmin(int32_t on);
/* test switch: FP64 recomputation of the winning closest-hit distance */
void orc_set_refine_hits(int32_t on);
float orc_rng(uint64_t seed, uint32_t pixel, uint32_t sample, uint32_t dim);

float orc_triangle_t(const float* a, const float* b, const float* c, const float* ray, float tmin, float tmax);
float orc_sphere_t(const float* center, float radius, const float* ray, float tmin, float tmax);
float orc_plane_t(const float* origin, const float* normal, const float* ray, float tmin, float tmax);
float orc_fresnel_dielectric(float cos_theta_i, float eta_i, float eta_t);
float orc_fresnel_conductor(float cos_theta_i, float eta, float k);
void orc_set_material_u(float u);  /* clear coat's layer draw in the BSDF probes (default 0.5) */
int32_t orc_bsdf_eval(const dfr_material* m, const float* wo, const float* wi, float* f, float* pdf);
int32_t orc_bsdf_sample(const dfr_material* m, const float* wo, float u, float v, float* wi, float* f,
                        float* pdf, int32_t* type);
int32_t orc_perez_sky(float turbidity, const float* sun, const float* dir, float* rgb);
int32_t orc_perez_sun(float turbidity, const float* sun, float* theta, float* spectrum, float* rgb);
int32_t orc_perez_pdf_many(float turbidity, const float* sun, int64_t n, const float* dirs, float* pdf);
int32_t orc_perez_sample_many(float turbidity, const float* sun, int64_t n, const float* uvw, const float* nrm,
                              float* out);
/* Image stage (AbstractImageKernel.java:1104-1324; GPURenderer.java:312-335):
 * orc_tone_map applies one operator (DFR_TONE_MAPPER_*) at exposure 1 to n
 * floats; orc_image_rgba8 turns n linear RGB means into RGBA8 as the
 * renderer does for `algorithm` (a DFR_ALGO_* ordinal). */
int32_t orc_tone_map(int32_t tone_mapper, int64_t n, const float* in, float* out);

/* The texture noise at n points (3 floats each): kind 0 Perlin noise, 1
 * Perlin turbulence over `octaves`, 2 simplex noise, 3 simplex fBm in [0, 1]
 * (frequency, gain, octaves). */
int32_t orc_noise(int32_t kind, int64_t n, const float* xyz, float frequency, float gain, int32_t octaves, float* out);
int32_t orc_image_rgba8(int32_t algorithm, int32_t tone_mapper, int64_t n, const float* rgb, uint8_t* rgba);

#ifdef __cplusplus
}
#endif

#endif
