/*
 * tc_oracle.h -- CPU oracle for the tessellated-cluster frame path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a serial C++ restatement of the
 * reference's GLSL stages, used by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py as the checker.  The product path never
 * links, loads or calls it.
 *
 * All tc_* records come from include/tc_abi.h.  Every uint64 "device
 * address" field of the records handed to the oracle holds a host pointer
 * (the oracle dereferences them), except the pure values the reference only
 * copies into records (template addresses, genClusterData): those are
 * carried through as numbers.
 *
 * Canonical order.  The reference appends to every list with atomicAdd, so
 * its order depends on scheduling.  The oracle defines the order a serial
 * execution yields when workgroups run one after another in index order:
 * (visible cluster, triangle) for classify, (split item, child) for split,
 * part index for instantiate, instance index for BLAS setup, CLAS index for
 * insertion.  The HIP kernels produce the same order.
 *
 * Floating point.  Built with -ffp-contract=off.  Evaluation order is the
 * left-to-right order written in the GLSL, with mat4*vec4 summed column by
 * column, length = sqrtf(dot), normalize = v * (1/length), GLSL round =
 * roundf (half away from zero).
 */
#ifndef TC_ORACLE_H
#define TC_ORACLE_H

#include "tc_abi.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- tessellation table (tessellation_table.cpp:36-100) ---- */
uint32_t oracle_table_num_vertices(void);
uint32_t oracle_table_num_triangles(void);
uint32_t oracle_table_num_configs(void);
const uint32_t* oracle_table_vertices(void);
const uint32_t* oracle_table_triangles(void);
const uint16_t* oracle_table_configs(void);
void oracle_table_lookup(tc_tess_table_entry out[4096]);

/* ---- tessellation.glsl utilities ---- */
uint32_t oracle_encode_barycentrics(const float wuv[3]);
void oracle_decode_barycentrics(uint32_t vtx, float wuv[3]);
void oracle_tess_factors(const float a[3], const float b[3], const float c[3], const float viewPos[3], float nearPlane,
                         float viewportY, float tessRate, float factors[3]);
void oracle_split_factor(const uint32_t f[3], uint32_t maxSplitFactor, uint32_t out[3]);
uint32_t oracle_get_config(const uint32_t factors[3], uint32_t vtxEncoded[3]);
void oracle_config_factors(uint32_t cfg, uint32_t out[3]);
void oracle_config_triangle_vertices(uint32_t cfg, uint32_t tri, uint32_t out[3]);
void oracle_config_vertex_barycentrics(uint32_t cfg, uint32_t vert, float wuv[3]);
float oracle_sample_height(const tc_displacement_texture* tex, float u, float v);
/* displacement.glsl:47-104: PN-triangle position of barycentrics (u, v, w)
 * over verts[0..2] / normals[0..2] (3 floats each) */
void oracle_pn(const float verts[9], const float normals[9], const float bary[3], float out[3]);
/* the ripple animation's sin / cos (FP64 reduction, FP32 Taylor series) */
float oracle_ripple_sin(float x);
float oracle_ripple_cos(float x);

/* ---- frame stages (host-address records) ---- */
typedef struct oracle_ctx {
  const tc_frame_constants* view;
  const tc_tessellation_table* tessTable; /* host addresses */
  tc_readback* readback;
  const tc_render_instance* instances; /* host addresses */
  tc_scene_building* build;            /* host addresses */
  const tc_displacement_texture* displacementTextures;
  uint32_t numRenderInstances;
  uint32_t numDisplacementTextures;
  /* value (not dereferenced) the genVertices field stands for in the
   * vertexBufferAddress records; 0 means use build->genVertices */
  uint64_t genVerticesValue;
  /* culling constants of the frame the HiZ was traced with (NULL: view)
   * and the far HiZ (host texels; NULL: no occlusion test) */
  const tc_frame_constants* viewLast;
  tc_hiz hiz;
} oracle_ctx;

/* nvhiz-update.comp.glsl: the far HiZ pyramid of a depth image (host) */
void oracle_hiz_build(const tc_hiz* hiz, const void* depth, uint32_t pixelStride);
/* culling.glsl:96-114 for one clip-space rectangle (xy clamped to [-1, 1]) */
int oracle_intersect_hiz(const tc_hiz* hiz, const tc_frame_constants* viewLast, const float clipMin[4],
                         const float clipMax[4]);

void oracle_instances_classify(const oracle_ctx* c, tc_config cfg);
void oracle_clusters_cull(const oracle_ctx* c, tc_config cfg);
void oracle_build_setup(const oracle_ctx* c, tc_config cfg, int mode);
void oracle_cluster_classify(const oracle_ctx* c, tc_config cfg);
void oracle_triangle_split(const oracle_ctx* c, tc_config cfg);
void oracle_triangle_tess_template_instantiate(const oracle_ctx* c, tc_config cfg);
void oracle_blas_setup_insertion(const oracle_ctx* c, tc_config cfg);
void oracle_blas_clusters_insert(const oracle_ctx* c, tc_config cfg, int transient);

void oracle_clas_instantiate(const oracle_ctx* c, tc_config cfg);

uint32_t oracle_num_split_passes(uint32_t splitFactor);

/* Closest-hit ray casting over every CLAS of every instance (after
 * oracle_blas_clusters_insert), no hierarchy: the reference's rgen / rchit /
 * shadow-ray semantics with a watertight triangle test.  rays == NULL casts
 * the camera's width x height primary rays.  flags: 1 cull back faces,
 * 2 shadow rays.  xforms: 24 floats per instance (objectToWorld,
 * worldToObject, 3x4 row-major). */
void oracle_trace(const oracle_ctx* c, const float* xforms, const tc_ray* rays, uint32_t numRays, uint32_t width,
                  uint32_t height, uint32_t flags, tc_hit* hits, tc_hit_decoded* decoded, uint32_t nthreads);

/* oracle_trace's camera rays for a subset of pixels (index y * width + x):
 * pixel-sampled parity at full viewport sizes. */
void oracle_trace_pixels(const oracle_ctx* c, const float* xforms, const uint32_t* pixels, uint32_t numPixels,
                         uint32_t width, uint32_t height, uint32_t flags, tc_hit* hits, tc_hit_decoded* decoded,
                         uint32_t nthreads);

/* render_raytrace.rgen.glsl:52-86: origin and direction of pixel (x, y). */
void oracle_primary_ray(const tc_frame_constants* view, uint32_t x, uint32_t y, uint32_t width, uint32_t height,
                        float origin[3], float direction[3]);

/* The chain of renderer_raytrace_clusters_tess.cpp:467-582 from
 * build_setup CLASSIFY through build_setup BUILD_BLAS (visible clusters and
 * instance states are inputs).  CLAS build and BLAS insertion follow as
 * separate calls. */
void oracle_frame_tessellate(const oracle_ctx* c, tc_config cfg, uint32_t numSplitPasses);

/* Our CLAS blocks from the transient build records (TC_FLAG_1X_TRANSIENT). */
void oracle_clas_build_transient(const oracle_ctx* c, tc_config cfg);

/* Same chain on nthreads threads for the CPU baseline (parallel over
 * clusters / items with a deterministic merge); identical output. */
void oracle_frame_tessellate_mt(const oracle_ctx* c, tc_config cfg, uint32_t numSplitPasses, uint32_t nthreads);

/* ---- depth adjoint checker (tc_oracle_adjoint.cpp; no reference
 * counterpart).  Height maps hold the sampled (normalised) texel values;
 * xforms holds 24 floats per instance (objectToWorld | worldToObject, 3x4
 * row-major, as tc_instance_xforms_host writes them). ---- */
typedef struct oracle_height_map {
  const float* texels; /* NULL: not differentiated */
  uint32_t width, height;
} oracle_height_map;

/* FP64 distance along each ray to the plane of its hit triangle with the
 * triangle's vertices re-displaced from the tape, texel perturbTexel of
 * map perturbMap offset by delta; -1 for rays without a hit. */
void oracle_depth_t64(const tc_ray* rays, const tc_hit* hits, const uint32_t* hitVertices, uint32_t numRays,
                      const tc_vertex_tape* tape, const oracle_height_map* maps, uint32_t numMaps, const float* xforms,
                      uint32_t perturbMap, uint64_t perturbTexel, double delta, double* t);

/* FP64 analytic adjoint: grads[map][texel] += gradT[r] * dt_r/dtexel. */
void oracle_depth_adjoint64(const tc_ray* rays, const tc_hit* hits, const uint32_t* hitVertices, uint32_t numRays,
                            const tc_vertex_tape* tape, const oracle_height_map* maps, uint32_t numMaps,
                            const float* xforms, const double* gradT, double* const* grads);

/* FP64 central differences (step eps) of
 *   L = scale * sum over rays with a hit and target >= 0 of (t - target)^2
 * with respect to the listed texels of one map. */
void oracle_depth_loss_fd64(const tc_ray* rays, const tc_hit* hits, const uint32_t* hitVertices, uint32_t numRays,
                            const tc_vertex_tape* tape, const oracle_height_map* maps, uint32_t numMaps,
                            const float* xforms, const float* target, double scale, uint32_t map, const uint64_t* texels,
                            uint32_t count, double eps, double* out);

#ifdef __cplusplus
}
#endif

#endif /* TC_ORACLE_H */
