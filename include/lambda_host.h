/* lambda_host.h -- C ABI of the host side above lambda_amd.h: the .nls
 * loader, tone mapping and PNG output, for callers that are not C++ (the
 * Python binding, tests).  C++ callers can use the Renderer class in
 * lambda_amd/csrc/host/renderer.hpp instead, which mirrors the reference's
 * include/renderer.hpp.
 */
#ifndef LAMBDA_HOST_H
#define LAMBDA_HOST_H

#include "lambda_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct lamh_scene lamh_scene;

/* nl::cg::load::loadNLS (lib/numli/cg.hpp:2455-2470).  Relative paths in
 * the scene (spectra CSV, OBJ meshes) resolve against the current working
 * directory, as in the reference.  data_dir (may be NULL) holds the CIE
 * tables.  Returns LAM_OK, LAM_E_IO or LAM_E_PARSE; err receives the text. */
int lamh_load_nls(const char *path, const char *data_dir, lamh_scene **out, char *err, size_t errlen);
int lamh_load_nls_text(const char *text, const char *data_dir, lamh_scene **out, char *err, size_t errlen);
void lamh_scene_free(lamh_scene *s);
/* Flattened description, valid until lamh_scene_free. */
const lam_scene_desc *lamh_scene_desc(const lamh_scene *s);
/* Hero wavelengths requested by the scene ("hero_samples", default 16). */
uint32_t lamh_scene_hero(const lamh_scene *s);
/* Names (for tests): object i / material i; NULL if out of range. */
const char *lamh_object_name(const lamh_scene *s, uint32_t i);
const char *lamh_material_name(const lamh_scene *s, uint32_t i);

/* Transpose of the loader's box filter coefficientλ(sampledλ)
 * (lib/numli/cg.hpp:252-272), which is linear in the sample values: for a
 * spectrum whose grid coefficients are mult * box_filter(samples), maps a
 * gradient with respect to the grid coefficients (dcoef, grid->n values,
 * e.g. a lam_adjoint albedo or film-IOR slot range) to the gradient with
 * respect to the CSV's value column (dsamples, n_samples values) at the
 * strictly increasing wavelengths sample_l.  LAM_E_INVALID on NULL
 * pointers, n_samples == 0 or unsorted wavelengths. */
int lamh_spectrum_grad_to_samples(const lam_grid *grid, const float *sample_l, size_t n_samples, float mult,
                                  const float *dcoef, float *dsamples);

/* Renderer::toneMap (src/renderer.cpp:149-176): log-average exposure
 * Y_MID/Ȳ then the koiFilmic curve per channel.  rgb_in/out: w*h*3. */
int lamh_tone_map(const float *rgb_in, uint32_t w, uint32_t h, float y_mid, float *rgb_out);
/* rendering::rgb24 + lodepng::encode (include/renderer.hpp:20-29,
 * src/renderer.cpp:187-190): sRGB OETF, 8-bit quantisation, RGB8 PNG. */
int lamh_write_png(const char *path, const float *rgb, uint32_t w, uint32_t h);
/* The same PNG from 8-bit RGB (w*h*3 bytes, e.g. lam_tone_map's output). */
int lamh_write_png8(const char *path, const uint8_t *rgb8, uint32_t w, uint32_t h);
/* Renderer::saveImage file name rule (src/renderer.cpp:182-185). */
int lamh_image_name(const char *scene_path, uint32_t spp, uint32_t bounces, float p, const char *suffix,
                    char *out, size_t outlen);

/* BVH8 statistics without a GPU (tests): nodes, leaves, depth, stack
 * bound, and the conservativeness of every quantized child box (returns 1
 * in *conservative when every decoded box contains its exact box). */
int lamh_bvh8_check(const float *tri_v, uint32_t n_tris, lam_scene_stats *stats, int *conservative);

/* The extend / shadow kernels' traversal on the host: builds the BVH8 as
 * lam_scene_upload does and traverses its 128-byte records with the
 * kernels' float operations (closest hit, or any hit when `any`), one ray
 * per query as intersect::scene (lib/numli/cg.hpp:1469-1489).  hits (closest,
 * may be NULL), occluded (may be NULL) as lam_extend / lam_occluded;
 * counts (may be NULL): 2 per ray, interior nodes fetched and triangles
 * tested — the independent count behind the extend roofline's bytes.
 * LAM_E_BVH if a ray needs more than 128 stack entries. */
int lamh_trace_counts(const lam_scene_desc *desc, const lam_ray *rays, size_t n, int any, lam_hit *hits,
                      uint32_t *occluded, uint32_t *counts);

#ifdef __cplusplus
}
#endif
#endif /* LAMBDA_HOST_H */
