/* CPU oracle for raw-native's ray-traced AO path: a plain C++ restatement of the
 * reference's float32 arithmetic, used ONLY as the checker (tests, smoke, the
 * cpu_baseline leg of bench.py). Nothing in the product links or loads it.
 *
 * Every function names the reference file:line it restates. All arrays are
 * row-major, y down, 3 floats per Vec3 (AoS), as raw::Buffer<Vec3> stores them.
 */
#ifndef RAW_AO_ORACLE_API_H
#define RAW_AO_ORACLE_API_H
#include <stddef.h>
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* hash01 (src/renderer/ray_ao.cpp:8-12, src/renderer/ssao.cpp:6-10). */
float orc_hash01(uint32_t x, uint32_t y, uint32_t s);

/* The cosine-weighted direction of sample s at pixel (x,y) for normal n
 * (src/renderer/ray_ao.cpp:14-17,32-38); writes dir[3] and origin[3]. */
void orc_ao_ray(uint32_t x, uint32_t y, uint32_t s, const float* p, const float* n,
                float* origin, float* dir);

/* intersectTri (raw/math/primitives.hpp:31-64); returns 1 on a hit, fills t,u,v. */
int orc_intersect(const float* ray_o, const float* ray_d, const float* tri9,
                  float* t, float* u, float* v);

/* Rasterize indexed triangle meshes into a G-buffer (src/renderer/raster.cpp:92-126,
 * default RasterOptions). Camera as raw::Camera (raw/scene/scene.hpp:15-20).
 * Outputs: depth w*h, position w*h*3, normal w*h*3, mask w*h (0/1). */
void orc_rasterize(const float* vpos, const float* vnrm, const int32_t* idx, size_t ntri,
                   const float eye[3], const float center[3], const float up[3],
                   float fovy, float aspect, float nearZ, float farZ, int w, int h,
                   float* depth, float* position, float* normal, uint8_t* mask);

/* The same over a scene of meshes concatenated in order (src/renderer/raster.cpp:107-124):
 * mesh m holds triangles below mesh_end[m]; albedo_out (w*h*3) gets the albedo of the mesh
 * whose fragment a pixel keeps (raster.cpp:82), 0 where nothing covers it. */
void orc_rasterize_scene(const float* vpos, const float* vnrm, const int32_t* idx, size_t ntri,
                         const uint32_t* mesh_end, const float* albedo, size_t nmesh,
                         const float eye[3], const float center[3], const float up[3],
                         float fovy, float aspect, float nearZ, float farZ, int w, int h,
                         float* depth, float* position, float* normal, float* albedo_out, uint8_t* mask);
/* perspective * lookAt (raw/math/mat.hpp:7-38), row-major. */
void orc_view_proj(const float eye[3], const float center[3], const float up[3], float fovy,
                   float aspect, float nearZ, float farZ, float out16[16]);
/* shade/shadeHDR (model 1, plane_a = AO) and shadeV2/shadeHDRV2 (model 2, plane_a = sky,
 * plane_sun = sun) (src/renderer/composite.cpp:8-65): frame clamped to [0,1], hdr not;
 * ldir = the light's direction (not negated), li its intensity. */
void orc_shade(const float* normal, const float* albedo, const uint8_t* mask, int w, int h, int model,
               const float* plane_a, const float* plane_sun, const float ldir[3], float li,
               float* frame, float* hdr);
/* maxRadiance (src/renderer/composite.cpp:82-88). */
float orc_max_radiance(const float* hdr, size_t npix);
/* computeMotion (src/renderer/motion.cpp:7-31): returns the valid count. */
int32_t orc_motion(const float* position, const uint8_t* mask, int w, int h, const float cur16[16],
                   const float prev16[16], float* motion);

/* The 14-triangle test scene (src/scene/scene.cpp:18-32) as one vertex/index list
 * (plane mesh then box mesh). Returns triangle count; pass NULL to size. */
size_t orc_test_scene(float* vpos, float* vnrm, int32_t* idx);

/* Binned-SAH BVH (src/renderer/bvh.cpp:35-92) over tris (9 floats each). */
void* orc_bvh_build(const float* tris9, size_t ntri);
void orc_bvh_free(void* bvh);
size_t orc_bvh_node_count(void* bvh);
/* Bvh::occluded (src/renderer/bvh.cpp:119-138). */
int orc_bvh_occluded(void* bvh, const float* o, const float* d, float tmax);

/* computeRTAO (src/renderer/ray_ao.cpp:20-58) over a built BVH.
 * mode: 0 BVH, 1 scan of every triangle, 2 skew (query distance x0.99), as
 * RAW_NATIVE_AO_ORACLE selects in src/renderer/accel.cpp:6-11.
 * Rows y0..y1-1 only (the rest of ao untouched); threads interleave rows. */
void orc_rtao(void* bvh, const float* tris9, size_t ntri, int mode,
              const float* position, const float* normal, const uint8_t* mask,
              int w, int h, int y0, int y1, int samples, float radius, int threads,
              float* ao);
/* orc_rtao on rows y0, y0 + ystep, ... below y1 only (a bounded sample of a large frame). */
void orc_rtao_rows(void* bvh, const float* tris9, size_t ntri, int mode,
                   const float* position, const float* normal, const uint8_t* mask,
                   int w, int h, int y0, int y1, int ystep, int samples, float radius, int threads,
                   float* ao);

/* computeRTVisibility (src/renderer/ray_ao.cpp:18-63), lighting model 2's visibility: ao
 * exactly as computeRTAO; sky = the fraction of the samples whose ray is open within radius
 * and also hits nothing with tMax 1e30; sun = 1 when the ray from the AO origin towards
 * normalize(to_light) hits nothing with tMax 1e30, or when sun_forced; all 1 off the mask.
 * mode as orc_rtao. */
void orc_rtvis(void* bvh, const float* tris9, size_t ntri, int mode,
               const float* position, const float* normal, const uint8_t* mask,
               int w, int h, int samples, float radius, const float to_light[3], int sun_forced,
               int threads, float* ao, float* sky, float* sun);

/* computeSSAO (src/renderer/ssao.cpp:11-39). */
void orc_ssao(const float* position, const float* normal, const uint8_t* mask,
              int w, int h, int samples, float radius, int threads, float* ao);

/* reconcile (src/cert/reconcile.cpp:4-21). */
void orc_reconcile(const float* approx, const float* truth, const uint8_t* mask,
                   int w, int h, float tolerance, float* error_map,
                   float* rmse, float* max_error, int32_t* pixels, int32_t* within);

/* PLOC build (src/renderer/rt_bvh.cpp:54-114). nodes out: 2n-1 records of
 * {lo[3], hi[3]} in boxes6 and {left, right, tri} in links3. Returns the root
 * (-1 when empty); rounds out. */
int32_t orc_ploc_build(const float* tris9, size_t ntri, int radius,
                       float* boxes6, int32_t* links3, uint32_t* rounds);
/* The Morton keys and the sorted order of the build's first two steps
 * (src/renderer/rt_bvh.cpp:56-73): keys[i] of triangle i, order[p] = triangle at
 * sorted position p. Also returns the centroid box lo[3] and scale in g4. */
void orc_ploc_keys(const float* tris9, size_t ntri, uint32_t* keys, uint32_t* order, float* g4);
/* FNV-1a tree hash (src/renderer/gpu/rt_bvh_parity.cpp:50-58). */
uint64_t orc_tree_hash(const float* boxes6, const int32_t* links3, size_t nnodes);
/* SAH cost (src/renderer/rt_bvh.cpp:116-123) and the binned-SAH one
 * (src/renderer/rt_bvh_sah.cpp:64-72). */
double orc_sah_cost(const float* boxes6, size_t nnodes, int32_t root);
double orc_binned_sah_cost(const float* tris9, size_t ntri, size_t* node_count);

/* Closest hit by a scan of every triangle (t < tmax; ties to the lower index) and
 * any hit (t < tmax), for the traversal checks (src/renderer/gpu/rt_bvh_parity.cpp:19-34). */
void orc_scan_rays(const float* tris9, size_t ntri, const float* rays6, const float* tmax,
                   size_t nrays, int any, int threads, float* t_out, int32_t* tri_out);

/* The R2 check's rays (src/renderer/gpu/rt_bvh_parity.cpp:36-48): rays8 {o, d, tMax = 1e30,
 * 0}; eye/target/up NULL for the scene without a camera (soup). cam_first_u picks the order of
 * the two draws of a camera ray's jitter. */
void orc_r2_rays(const float* eye, const float* target, const float* up, size_t n, uint64_t seed,
                 int cam_first_u, float* rays8);
/* bruteForce (rt_bvh_parity.cpp:19-34): hits4 {t, u, v, tri (int32 bits)} of the nearest hit
 * (ties to the lower index; -1, 0, 0, -1 on a miss) and the second-nearest t. */
void orc_scan_hits(const float* tris9, size_t ntri, const float* rays8, size_t nrays, int threads,
                   float* hits4, float* t2);
/* rt::closest (src/renderer/rt_bvh.cpp:125-152) over a tree of boxes6/links3 from root. */
void orc_tree_closest(const float* boxes6, const int32_t* links3, int32_t root, const float* tris9,
                      const float* rays8, size_t nrays, int threads, float* hits4);

/* std::sin and std::cos of every sample angle 6.2831853f * k / 2^24, k < 2^24. */
void orc_libm_angles(float* s, float* c);

/* Owned scenes (src/renderer/rt_scenes.cpp, src/renderer/swr_scenes.cpp) as triangle
 * lists, 9 floats each: name in {"retro_room","iso_street","clip_stress","dense_block",
 * "soup"}; seed/count used by "soup" only. Returns the count; NULL out to size. */
size_t orc_scene_tris(const char* name, size_t count, uint64_t seed, float* tris9);

#ifdef __cplusplus
}
#endif
#endif
