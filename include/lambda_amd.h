/* lambda_amd.h -- C ABI of the MI355X render loop of lambda.
 *
 * This is the drop-in boundary under `Renderer::render`
 * (include/renderer.hpp:39 of the reference).  Host C++ (the .nls loader,
 * the Renderer class and the CLI, declared in lambda_host.h) flattens an
 * nl::cg::scene into a lam_scene_desc, uploads it once, and calls lam_render.
 * All pointers crossing this ABI are HOST pointers unless a function name
 * ends in `_dev`; device memory is owned by the context and never escapes.
 *
 * Error behaviour: every entry point returns LAM_OK (0) or a negative
 * LAM_E_* code and never throws; lam_last_error() gives the text.  The
 * reference's `render` is void and traps on FP errors (src/main.cpp:15);
 * here device code never traps, errors come back as codes.
 *
 * Threading: one lam_ctx is used by one host thread at a time; every call
 * is synchronous on return (the context owns one HIP stream).
 */
#ifndef LAMBDA_AMD_H
#define LAMBDA_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes --------------------------------------------------- */
#define LAM_OK 0
#define LAM_E_INVALID (-1)  /* bad argument / malformed scene description */
#define LAM_E_DEVICE (-2)   /* HIP runtime error (allocation, launch, fault) */
#define LAM_E_NOSCENE (-3)  /* call needs an uploaded scene */
#define LAM_E_LIMIT (-4)    /* a size exceeds a documented maximum */
#define LAM_E_IO (-5)       /* file could not be read or written */
#define LAM_E_PARSE (-6)    /* .nls / OBJ / CSV syntax or schema error */
#define LAM_E_BVH (-7)      /* BVH exceeds the traversal stack bound */

#define LAM_NONE 0xFFFFFFFFu /* "no object" / miss marker */

/* ---- enums (values mirror the reference's variant alternatives) ------ */
/* Object = std::variant<sphere, plane, trimesh>   (lib/numli/cg.hpp:857) */
enum { LAM_OBJ_SPHERE = 1, LAM_OBJ_PLANE = 2, LAM_OBJ_TRIMESH = 3 };
/* Material = variant<lambertian, blinn, thinfilm, microfacet, emitter,
 * diremitter>                                      (lib/numli/cg.hpp:1084) */
enum {
  LAM_MAT_NONE = 0,
  LAM_MAT_LAMBERTIAN = 1,
  LAM_MAT_BLINN = 2,
  LAM_MAT_THINFILM = 3,
  LAM_MAT_MICROFACET = 4,
  LAM_MAT_EMITTER = 5,
  LAM_MAT_DIREMITTER = 6
};
/* Light = variant<ambient, point, dir, sphere, quaddir> (cg.hpp:904) */
enum {
  LAM_LIGHT_AMBIENT = 0,
  LAM_LIGHT_POINT = 1,
  LAM_LIGHT_DIR = 2,
  LAM_LIGHT_SPHERE = 3,
  LAM_LIGHT_QUADDIR = 4
};

/* ---- scene description (flat, SoA where it is large) ---------------- */

/* Coefficient-spectrum grid: n values at inf + i*delta, delta =
 * (sup-inf)/(n-1) in float.  The reference fixes n=80 on [400,720]
 * (coefficientλ, cg.hpp:241-246); hero wavelengths are drawn on [inf,sup]
 * (sample::spectrum, cg.hpp:1517-1531). */
typedef struct {
  uint32_t n;
  float inf, sup;
} lam_grid;

/* Rows 0..2 of a 4x4 affine matrix, row-major (transform, cg.hpp:422-487).
 * Row 3 is implicitly (0,0,0,1). */
typedef struct {
  float M[12];
  float Minv[12];
} lam_xform;

typedef struct {
  uint32_t type;     /* LAM_OBJ_* */
  uint32_t material; /* index into materials */
  uint32_t mesh;     /* trimesh: mesh id (informational) */
  uint32_t pad;
  lam_xform T;
} lam_object;

/* Spectrum slots (indices into the spectra table, -1 = unused):
 *   lambertian: [0]=albedo                       (cg.hpp:921-926)
 *   blinn:      [0]=Kd [1]=Ks [2]=Kt [3]=Le [4]=ior (cg.hpp:927-993)
 *   thinfilm:   [0]=film_ior [1]=base_ior         (cg.hpp:995-1052)
 *   emitter / diremitter: [0]=radiance            (cg.hpp:1069-1081)   */
typedef struct {
  uint32_t type;
  int32_t spec[5];
  float alpha;    /* blinn / thinfilm glossiness α */
  float d;        /* thinfilm thickness in nm */
  float p_d, p_spec, w_r, w_t; /* blinn::init() results (cg.hpp:931-940) */
  float specular; /* diremitter exponent */
  float pad;
} lam_material;

typedef struct {
  uint32_t type;   /* LAM_LIGHT_* */
  uint32_t object; /* sphere / quaddir: the emitting object; else LAM_NONE */
  int32_t spec;    /* radiance or radiant intensity spectrum */
  float size;      /* sphere light radius (cg.hpp:2121) */
  float inv_area;  /* quaddir 1/(|col0||col1|) (cg.hpp:2098-2099) */
  float specular;  /* quaddir exponent (cg.hpp:2087-2088) */
  float pos[3];    /* point light position */
  float dir[3];    /* directional light direction (as loaded) */
} lam_light;

/* camera after camera::init (cg.hpp:1151-1165, 2039-2063) */
typedef struct {
  float pos[3], x[3], y[3], z[3];
  float D, delta, w, h, dof, fov;
  uint32_t width, height;
} lam_camera;

typedef struct {
  lam_grid grid;
  uint32_t n_spectra;
  const float *spectra; /* n_spectra * grid.n coefficients */
  const float *cmf;     /* 3 * grid.n: CIE 1931 x̄,ȳ,z̄ box-filtered onto grid */
  uint32_t n_objects;
  const lam_object *objects;
  uint32_t n_materials;
  const lam_material *materials;
  uint32_t n_lights;
  const lam_light *lights;
  /* every trimesh instance flattened to world space, in object order */
  uint32_t n_tris;
  const float *tri_v;       /* 9 per tri: v0,v1,v2 world space            */
  const float *tri_n;       /* 9 per tri: vertex normals, normalToWorld()   */
  const float *tri_gn;      /* 3 per tri: geometric normal, world, unit     */
  const uint32_t *tri_obj;  /* object index of the instance                */
  const uint32_t *tri_prim; /* face index inside the instanced mesh        */
  lam_camera camera;
} lam_scene_desc;

/* ---- ray batches for the extend / shadow test hooks ------------------ */
/* 32 B.  A hit is accepted for t in [BIAS, tmax]; `skip` is an object id
 * to ignore (LAM_NONE for none), as intersect::scene's `skip`
 * (cg.hpp:1469-1489). */
typedef struct {
  float o[3];
  float tmax;
  float d[3];
  uint32_t skip;
} lam_ray;

/* 16 B.  Closest hit; ties in t go to the lowest (obj, prim).  Miss:
 * obj = LAM_NONE, t = FLT_MAX.  `tri` is the index into the desc's flat
 * triangle arrays (LAM_NONE for spheres/planes). */
typedef struct {
  float t;
  uint32_t tri;
  uint32_t obj;
  uint32_t prim;
} lam_hit;

/* ---- render parameters ---------------------------------------------- */
#define LAM_FLAG_MEGAKERNEL 0x1u /* one thread per path instead of queues */
#define LAM_FLAG_NO_GRAPH 0x2u   /* launch the bounce loop without a hipGraph */
#define LAM_FLAG_PROFILE 0x4u    /* time every kernel with HIP events */
#define LAM_FLAG_NO_TAIL 0x8u    /* keep short queues in the wavefront (no finish kernel) */
#define LAM_FLAG_NO_PIPE 0x10u   /* trace passes one after another on one stream (no overlap) */
#define LAM_FLAG_ACCUMULATE 0x20u /* add the passes to the film of the previous render (progressive) */
#define LAM_FLAG_PROFILE_FETCH 0x40u /* with LAM_FLAG_PROFILE: the extend and shadow kernels count the
                                        BVH nodes and triangles they fetch (lam_timing) */

/* lam_scene_upload_ex flags */
#define LAM_UPLOAD_GPU_BVH 0x1u /* build the BVH8 on the device (linear BVH; faster build, slower trace) */

typedef struct {
  uint32_t spp;             /* Renderer::SPP            (renderer.hpp:36) */
  uint32_t max_scatterings; /* Renderer::MAX_SCATTERINGS (renderer.hpp:35) */
  float sample_p;           /* Renderer::SAMPLE_P       (renderer.hpp:37) */
  uint32_t hero;            /* hero wavelengths per path: 16 (reference) or 4 */
  uint64_t seed;            /* counter-RNG key */
  uint32_t spp_offset;      /* first sample index (progressive / resume) */
  uint32_t flags;           /* LAM_FLAG_* */
} lam_render_params;

typedef struct {
  uint32_t n_nodes, n_leaves, n_tris, depth;
  uint32_t stack_bound; /* worst-case traversal stack entries */
  uint32_t n_analytic;  /* spheres + planes tested outside the BVH */
  uint64_t device_bytes;
  double sah_cost;
} lam_scene_stats;

/* Kernel timing accumulated since the last reset (LAM_FLAG_PROFILE, or the
 * extend/occlusion hooks which are always timed).  extend_rays and
 * shadow_rays also count the rays the passes of lam_render and
 * lam_render_guided trace (not those of LAM_FLAG_MEGAKERNEL). */
typedef struct {
  double extend_ms, shadow_ms, shade_ms, raygen_ms, film_ms, other_ms;
  uint64_t extend_launches, shadow_launches, shade_launches;
  uint64_t extend_rays, shadow_rays;
  double render_ms; /* wall time of the last lam_render on the stream */
  /* lam_render_guided: training of the last call */
  uint64_t guide_records;  /* training records written (all training passes) */
  double guide_train_ms;   /* training passes + Adam steps, on the stream */
  uint64_t guide_dropped;  /* continuations dropped: weight not finite */
  uint64_t guide_zero_pdf; /* guide samples whose mixture density was zero */
  double guide_loss_first, guide_loss_last; /* training loss, first and last step */
  /* LAM_FLAG_PROFILE | LAM_FLAG_PROFILE_FETCH renders: interior nodes
   * fetched, triangles tested and rays traced by the extend and the shadow
   * kernels (the finish kernel's rays are not among them) */
  uint64_t extend_nodes, extend_tris, extend_traced;
  uint64_t shadow_nodes, shadow_tris, shadow_traced;
  /* the wavefront passes of the last lam_render / lam_render_guided (its
   * first chunk, after any training passes): samples of every pixel per
   * pass and the number of passes */
  uint32_t pass_samples, passes;
} lam_timing;

/* ---- context ---------------------------------------------------------- */
typedef struct lam_ctx lam_ctx;

int lam_create(lam_ctx **out, int device);
void lam_destroy(lam_ctx *ctx);
const char *lam_last_error(const lam_ctx *ctx);
/* Library version string. */
const char *lam_version(void);

/* Copies the scene to HBM: quantized BVH8 (128-B nodes), 48-B triangle
 * records, shading records, material/spectra tables.  Replaces the scene
 * build of loadTriMesh/buildBVH (cg.hpp:732-749, 2287-2304) and the
 * per-render read of nl::cg::scene (cg.hpp:1167-1180). */
int lam_scene_upload(lam_ctx *ctx, const lam_scene_desc *desc);
/* lam_scene_upload with LAM_UPLOAD_* flags.  LAM_UPLOAD_GPU_BVH builds the
 * BVH8 on the device: a linear BVH over 30-bit Morton codes (radix sort),
 * collapsed and quantized as the host builder does; milliseconds instead of
 * seconds for 10 M triangles, a somewhat slower trace, the same images (no
 * counterpart in the reference, whose bvh<>::build runs per mesh on the
 * host, cg.hpp:632-749). */
int lam_scene_upload_ex(lam_ctx *ctx, const lam_scene_desc *desc, uint32_t flags);
int lam_get_scene_stats(const lam_ctx *ctx, lam_scene_stats *out);

/* Renderer::render (src/renderer.cpp:14-61).  linrgb: w*h*3 floats, top
 * row first (the reference's (h-i-1)*w+j flip).  xyz (optional, may be
 * NULL): w*h*3 floats, compensated per-pixel XYZ sums before the
 * 1/(SPP*hero) normalisation, same orientation. */
int lam_render(lam_ctx *ctx, const lam_render_params *p, float *linrgb,
               float *xyz);

/* Renderer::tracePath (src/renderer.cpp:64-146) for a batch of
 * (pixel, sample) pairs: pixel = i*w+j with row i counted from the bottom,
 * as in render.  out: n*hero floats, the spectral radiance Li of each path
 * (before the 1/prob_λ film weight). */
int lam_trace_paths(lam_ctx *ctx, const lam_render_params *p,
                    const uint32_t *pixel, const uint32_t *sample, size_t n,
                    float *radiance_out);

/* ---- differentiation (path-replay adjoint) ---------------------------- */
/* Which material parameters are differentiated, and where their gradients
 * go in the output vector.  Per material (n_materials entries each; -1 =
 * not differentiated):
 *   mat_d       thin-film thickness d (thinfilm::d, cg.hpp:998)      1 slot
 *   mat_ior     thin-film film IOR coefficients (film_ior, cg.hpp:997) grid.n slots
 *   mat_albedo  Lambertian albedo coefficients (albedo, cg.hpp:923)    grid.n slots
 * Gradients with respect to a material's coefficient spectrum are per grid
 * coefficient; a flat spectrum's scalar gradient is their sum. */
typedef struct {
  uint32_t n_params;
  uint32_t n_materials; /* entries of each map: the uploaded scene's material count */
  const int32_t *mat_d;
  const int32_t *mat_ior;
  const int32_t *mat_albedo;
} lam_diff_params;

/* Gradient of an image loss with respect to the parameters above:
 * grad[j] = Σ_pixels Σ_c dloss_drgb[px][c] · ∂rgb[px][c]/∂θ_j for the image
 * lam_render(p) produces (same seed, spp, spp_offset).  dloss_drgb: w*h*3,
 * top row first (as lam_render's output).  Each path is replayed from its
 * counter-RNG key, so device memory does not grow with path length.
 * Sampling decisions are held fixed (none of these parameters changes a
 * sampled direction, a pdf or a Russian-roulette decision). */
int lam_adjoint(lam_ctx *ctx, const lam_render_params *p, const lam_diff_params *dp,
                const float *dloss_drgb, float *grad);

/* ---- learned path guiding -------------------------------------------- */
/* No reference counterpart: a proposal distribution for the continuation
 * direction of tracePath's BSDF sample (src/renderer.cpp:116-145), learned
 * online.  A small MLP (6 inputs: position in the scene box, the outgoing
 * direction as two angles, hero wavelength, each in [0,1]; 12-frequency
 * encoding; 4 hidden layers of 64; bf16 weights on MFMA with fp32
 * accumulation) outputs 8 von Mises-Fisher lobes (40 raw values: per lobe
 * an unnormalised mean direction, log κ and a weight logit).  Where the
 * reference's sampler picks a lobe named in `lobes` (a Lambertian cosine
 * lobe, the Kd lobe or the specular reflection lobe of Blinn), the
 * direction is drawn from the mixture with probability alpha and from the
 * lobe's own distribution otherwise, weighted by one-sample MIS over both,
 * so the image converges to lam_render's. */
typedef struct {
  float alpha;             /* guide-branch probability, [0, 1); 0 = lam_render */
  uint32_t train_passes;   /* the first train_passes samples per pixel train the network online, traced in
                              passes of m = min(8, train_passes / 4) samples (at least 1) */
  uint32_t steps_per_pass; /* Adam steps per training sample: m x steps_per_pass after each training pass */
  uint32_t batch;          /* training records per step, multiple of 256, <= 65536 */
  float lr;                /* Adam step size */
  uint64_t seed;           /* network initialisation (if none yet) and minibatches */
  uint32_t materials;      /* material types guided: bit (1 << LAM_MAT_*); 0 = all supported */
  uint32_t max_vertex;     /* guide path vertices of bounce index < max_vertex (0 = every vertex) */
  uint32_t lobes;          /* LAM_GUIDE_* lobes proposed; 0 = LAM_GUIDE_LAMBERTIAN | LAM_GUIDE_BLINN_KD */
} lam_guide_params;
/* lam_guide_params.lobes: the sampler lobes the mixture stands in for */
#define LAM_GUIDE_LAMBERTIAN 0x1u /* a Lambertian material's cosine lobe (cg.hpp:1755-1778) */
#define LAM_GUIDE_BLINN_KD 0x2u   /* the Kd lobe of a Blinn material (cg.hpp:1853-1861) */
#define LAM_GUIDE_BLINN_SPEC 0x4u /* the specular reflection lobe of a Blinn material without
                                     transmission (Kt = 0; cg.hpp:1815-1847) */

/* lam_render with the learned proposal.  The network persists in the
 * context (initialised from g->seed on first use, or by lam_guide_reset /
 * lam_guide_set_params) and keeps training across calls. */
int lam_render_guided(lam_ctx *ctx, const lam_render_params *p, const lam_guide_params *g,
                      float *linrgb, float *xyz);
/* number of fp32 network parameters (W1[64][150] b1 W2[64][64] b2 W3 b3 W4
 * b4 W5[40][64] b5; feature f < 144 is sin/cos(π 2^j x_d) with f = 24d+2j+c,
 * 144+d is x_d) */
size_t lam_guide_param_count(void);
int lam_guide_reset(lam_ctx *ctx, uint64_t seed);
int lam_guide_set_params(lam_ctx *ctx, const float *theta);
int lam_guide_get_params(lam_ctx *ctx, float *theta);
/* Network outputs (n x 40) for inputs (n x 6), through the MFMA kernel. */
int lam_guide_infer(lam_ctx *ctx, const float *in, size_t n, float *raw);
/* Training loss -Σ (w_i/Σw) log p(ω_i | in_i) and its gradient (param_count
 * floats) for rows (n x 10: in[6], ω[3], w), 1 <= n <= 65536; raw
 * (optional, n x 40): the training forward's network outputs. */
int lam_guide_loss_grad(lam_ctx *ctx, const float *rows, size_t n, float *loss, float *grad, float *raw);
/* `steps` Adam steps (β1 0.9, β2 0.99) on one batch of rows; *loss = the
 * last step's loss before its update. */
int lam_guide_train(lam_ctx *ctx, const float *rows, size_t n, uint32_t steps, float lr, float *loss);

/* Closest-hit query, intersect::scene (cg.hpp:1469-1489). */
int lam_extend(lam_ctx *ctx, const lam_ray *rays, size_t n, lam_hit *hits);
/* Shadow query: occluded[i] = 1 iff a hit exists in [BIAS, tmax]
 * excluding rays[i].skip (the shadow calls in cg.hpp:1596-1698). */
int lam_occluded(lam_ctx *ctx, const lam_ray *rays, size_t n,
                 uint32_t *occluded);
/* Device-pointer variants for rays already resident in HBM; `counts`
 * (optional, device, 2*n uint32) receives interior nodes fetched and
 * triangles tested per ray. */
int lam_extend_dev(lam_ctx *ctx, const lam_ray *d_rays, size_t n,
                   lam_hit *d_hits, uint32_t *d_counts);
/* Device-pointer variants of lam_occluded (measurement: the shadow
 * kernel's roofline); `counts` (optional, device, 2*n uint32) receives
 * interior nodes fetched and triangles tested per ray, and the timed form
 * runs `reps` (<= 4096) back-to-back launches between one pair of events. */
int lam_occluded_dev(lam_ctx *ctx, const lam_ray *d_rays, size_t n,
                     uint32_t *d_occluded, uint32_t *d_counts);
int lam_occluded_dev_timed(lam_ctx *ctx, const lam_ray *d_rays, size_t n,
                           uint32_t *d_occluded, uint32_t reps, double *avg_ms);
/* lam_extend_dev for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  No counterpart in the reference.  The caller passes the
 * device rays and hits; the ray-queue counter, the overflow flag and the
 * traversal-stack spill columns live in a static device array of
 * 54,526,208 bytes that every GPU the library runs on holds until the
 * process ends (a grid of at most 2,048 64-lane blocks).  Results are those
 * of lam_extend_dev bit for bit.  n <= 2^31 (LAM_E_LIMIT beyond).  Calls
 * that overlap on one GPU (from several contexts or threads) race on that
 * array.  Synchronous on return, like lam_extend_dev. */
int lam_extend_dev_noalloc(lam_ctx *ctx, const lam_ray *d_rays, size_t n,
                           lam_hit *d_hits);
/* Measurement hook: `reps` (<= 4096) back-to-back closest-hit launches
 * over device rays between one pair of HIP events on the context's stream;
 * *avg_ms receives the mean duration of one launch. */
int lam_extend_dev_timed(lam_ctx *ctx, const lam_ray *d_rays, size_t n,
                         lam_hit *d_hits, uint32_t reps, double *avg_ms);

/* Test hook: the device's elementary functions on n inputs (fn 0 sinf,
 * 1 cosf, 2 powf(a,b), 3 atan2f(a,b), 4 sqrtf(a), 5 a/b), for bitwise
 * comparison with the CPU oracle. */
int lam_debug_math(lam_ctx *ctx, int fn, const float *a, const float *b, size_t n, float *out);

/* Test hook (no counterpart in the reference): the device BVH builder's
 * stable radix sort of (key, value) pairs by the low `bits` (1..32) bits of
 * the keys, on host arrays of n <= 2^28 pairs; keys_out/vals_out receive the
 * sorted pairs. */
int lam_debug_radix_sort(lam_ctx *ctx, const uint32_t *keys, const uint32_t *vals, size_t n, uint32_t bits,
                         uint32_t *keys_out, uint32_t *vals_out);

/* ---- progressive rendering and checkpoints ------------------------------
 * No counterpart in the reference (progress/ETA output is listed as future
 * work in its README.md:24).  A render with LAM_FLAG_ACCUMULATE adds passes
 * spp_offset .. spp_offset+spp-1 to the film of the previous render (same
 * camera and hero count) and returns the image of everything accumulated;
 * each pixel still receives its samples in sample order, so rendering in
 * pieces (or from a restored checkpoint) gives the bits of one render.
 *
 * lam_film_get copies the compensated XYZ film (width*height*8 floats per
 * pixel: X, Y, Z sums, 0, X, Y, Z compensations, 0) and the samples per
 * pixel and hero count it holds; lam_film_set restores such a checkpoint
 * (for the current camera). */
int lam_film_get(lam_ctx *ctx, float *film, uint32_t *spp, uint32_t *hero);
/* Tone map + 8-bit sRGB of the last rendered image on the device
 * (Renderer::toneMap, src/renderer.cpp:149-176: log-average exposure
 * y_mid/Ȳ then the koiFilmic curve, cg.hpp:181-194; rendering::rgb24,
 * include/renderer.hpp:20-29: sRGB OETF and 8-bit rounding).  rgb8:
 * width*height*3 bytes, top row first, ready for lamh_write_png8.  The
 * log-average is summed in double on the device, so Ȳ may differ from the
 * host's float sum (lamh_tone_map) in its last bits. */
int lam_tone_map(lam_ctx *ctx, float y_mid, uint8_t *rgb8);
int lam_film_set(lam_ctx *ctx, const float *film, uint32_t spp, uint32_t hero);

/* Replace the uploaded scene's camera (another view or resolution of the
 * same geometry without rebuilding the BVH).  No counterpart in the
 * reference, whose camera is part of the loaded scene (cg.hpp:2039-2063). */
int lam_set_camera(lam_ctx *ctx, const lam_camera *cam);

int lam_get_timing(const lam_ctx *ctx, lam_timing *out);
int lam_reset_timing(lam_ctx *ctx);

/* ---- wavefront launch configuration ------------------------------------
 * No counterpart in the reference (its render loop is one thread per
 * pixel, src/renderer.cpp:35-62).  The defaults suit frames of 10^5..10^7
 * pixels; images do not depend on these values (same paths, same film
 * order), only the time does. */
typedef struct {
  uint32_t lanes;          /* passes in flight, each on its own stream: 1..4 (default 3) */
  uint32_t tail_threshold; /* the finish kernel takes over a queue at or below this length (default 131072) */
  uint32_t tail_div;       /* ... capped at 1/tail_div of the chunk's paths, >= 1 (default 16) */
  uint32_t max_iters;      /* synchronised bounces before the finish kernel takes the rest, >= 1 (default 12) */
  uint32_t trace_per_cu;   /* extend / shadow blocks (64 lanes) per CU in render passes: 1..32 (default 16) */
  uint32_t batch_per_cu;   /* ... for ray batches (lam_extend, lam_occluded, lam_extend_dev): 1..32 (default 24) */
  uint32_t finish_per_cu;  /* finish-kernel blocks (64 lanes) per CU: 1..32 (default 10) */
  uint32_t shade_per_cu;   /* shade-kernel blocks (256 lanes) per CU: 1..8 (default 3) */
  uint32_t samples_per_pass; /* at most this many samples of every pixel one wavefront pass traces together:
                                1..32 (default 32), at most 64 Mi paths per pass (frames of at most 64 Mi pixels);
                                a render of spp samples takes ceil(spp / g) per pass, g >= 2 passes rounded up to
                                a multiple of the lanes, so that every lane traces beside the others; the film still
                                adds each pixel's samples in sample order */
} lam_wave_config;
int lam_get_wave_config(const lam_ctx *ctx, lam_wave_config *out);
int lam_set_wave_config(lam_ctx *ctx, const lam_wave_config *cfg);

#ifdef __cplusplus
}
#endif
#endif /* LAMBDA_AMD_H */
