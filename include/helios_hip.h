/*
 * helios_hip.h — C ABI of the MI355X (gfx950) path-tracing hot path.
 *
 * This library replaces Helios's Vulkan ray-tracing render path:
 *   - helios::PathIntegrator            (include/gfx/path_integrator.h:9-62,
 *                                        src/engine/gfx/path_integrator.cpp:48-200)
 *   - the GPU packing of helios::Scene   (src/engine/resource/scene.cpp:915-1311)
 *   - BLAS per mesh / TLAS build         (src/engine/resource/mesh.cpp:46-117,
 *                                        src/engine/gfx/renderer.cpp:113-181)
 *   - Hosek-Wilkie sky bake              (src/engine/gfx/hosek_wilkie_sky_model.cpp:658-764)
 *   - output ping-pong / clear / tonemap (src/engine/gfx/renderer.cpp:183-223,369-430)
 *
 * Conventions
 *   - Every entry point returns an hh_status; nothing throws across the ABI.
 *     hh_last_error() returns a thread-local message for the last failure.
 *   - Handles are opaque; inputs are copied at call time (the caller keeps its
 *     host arrays); the library owns all device memory.
 *   - A context is externally synchronized: one host thread at a time.
 *   - Render calls are asynchronous on the context's HIP stream; hh_sync waits.
 *     A frame may run part of its work on a second stream of the context; that
 *     work is joined back into the context stream before the frame's last
 *     launch, so a wait on the context stream covers the whole frame.
 *   - Surfaces are RGBA32F, row-major, row 0 = NDC y = -1 (the reference's
 *     storage-image orientation, renderer.cpp:396-399 flips it for display).
 */
#ifndef HELIOS_HIP_H
#define HELIOS_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HH_ABI_VERSION 1

/* Scene limits of the reference (include/resource/scene.h:14-17). */
#define HH_MAX_SCENE_MESH_INSTANCE_COUNT 1024
#define HH_MAX_SCENE_LIGHT_COUNT 100000
#define HH_MAX_SCENE_MATERIAL_COUNT 4096
#define HH_MAX_SCENE_MATERIAL_TEXTURE_COUNT (HH_MAX_SCENE_MATERIAL_COUNT * 4)

typedef enum hh_status {
    HH_OK = 0,
    HH_ERR_INVALID_ARG = 1,
    HH_ERR_OUT_OF_MEMORY = 2,
    HH_ERR_LIMIT_EXCEEDED = 3,
    HH_ERR_DEVICE = 4,
    HH_ERR_NAN_DETECTED = 5,
    HH_ERR_NOT_READY = 6,
    HH_ERR_IO = 7              /* a file could not be read, written or parsed */
} hh_status;

typedef struct hh_context_t* hh_context;
typedef struct hh_scene_t* hh_scene;
typedef struct hh_surface_t* hh_surface;
typedef struct hh_integrator_t* hh_integrator;
typedef struct hh_denoiser_t* hh_denoiser;
typedef struct hh_debug_rays_t* hh_debug_rays;

/* ---------------------------------------------------------------------------
 * Data layouts shared with the reference (byte compatible).
 * ------------------------------------------------------------------------- */

/* helios::Vertex, include/resource/mesh.h:10-17 (5 x vec4 = 80 B).
 * position.w carries the submesh index as the reference loader writes it
 * (src/engine/core/resource_manager.cpp:467-473). */
typedef struct hh_vertex {
    float position[4];
    float tex_coord[4];
    float normal[4];
    float tangent[4];
    float bitangent[4];
} hh_vertex;

/* helios::SubMesh, include/resource/mesh.h:19-29 (numeric fields only). */
typedef struct hh_submesh {
    uint32_t mat_idx;      /* index into the mesh's material list */
    uint32_t index_count;
    uint32_t vertex_count;
    uint32_t base_vertex;
    uint32_t base_index;
} hh_submesh;

/* helios::MaterialType, include/resource/material.h:12-16. */
typedef enum hh_material_type { HH_MATERIAL_OPAQUE = 0, HH_MATERIAL_TRANSPARENT = 1 } hh_material_type;

/* helios::Material (include/resource/material.h:25-99): host-side description.
 * Texture ids come from hh_scene_add_texture, -1 = none. The packing into the
 * 80-B GPU MaterialData record (src/engine/resource/scene.cpp:25-32,970-1109)
 * happens at commit. */
typedef struct hh_material_desc {
    float albedo_value[4];    /* sRGB-ish; packed as pow(rgb, 2.2) (scene.cpp:1017) */
    float emissive_value[4];
    float metallic_value;
    float roughness_value;
    int32_t albedo_texture;
    int32_t normal_texture;
    int32_t roughness_texture;
    int32_t roughness_channel;  /* roughness_texture_info().array_index, scene.cpp:1055 */
    int32_t metallic_texture;
    int32_t metallic_channel;   /* metallic_texture_info().array_index, scene.cpp:1076 */
    int32_t emissive_texture;
    uint32_t type;              /* hh_material_type */
    uint32_t alpha_test;        /* is_alpha_tested() */
} hh_material_desc;

typedef enum hh_texture_format {
    HH_TEXTURE_RGBA8_UNORM = 0,
    HH_TEXTURE_RGBA8_SRGB = 1,  /* rgb decoded sRGB->linear per texel, alpha linear */
    HH_TEXTURE_RGBA32F = 2
} hh_texture_format;

/* helios::MeshNode (include/resource/scene.h:143-160): one TLAS instance.
 * model = global_transform() (column-major 4x4, glm layout); normal =
 * normal_matrix() = global_transform_without_scale() (scene.cpp:277-280). */
typedef struct hh_instance_desc {
    float model[16];
    float normal[16];
    uint32_t mesh;               /* id from hh_scene_add_mesh */
    int32_t material_override;   /* material id or -1 */
} hh_instance_desc;

/* Light node kinds; numeric values are the GPU light types (common.glsl:14-18). */
typedef enum hh_light_type {
    HH_LIGHT_DIRECTIONAL = 0,
    HH_LIGHT_SPOT = 1,
    HH_LIGHT_POINT = 2
} hh_light_type;

/* DirectionalLightNode / SpotLightNode / PointLightNode (scene.h:162-245).
 * forward = orientation * +Z (scene.cpp:203-208). Defaults of the reference:
 * radius 0.1 (directional) / 5 (spot, point); cone angles 40/50 degrees. */
typedef struct hh_light_desc {
    uint32_t type;          /* hh_light_type */
    float color[3];
    float intensity;
    float radius;
    float position[3];      /* global_position() for spot/point */
    float forward[3];       /* forward() for directional/spot */
    float inner_cone_deg;
    float outer_cone_deg;
} hh_light_desc;

/* Hosek-Wilkie sky parameters. The reference hard-codes turbidity 4,
 * ground albedo 0.1 and a sun luminance normalized to 1.15
 * (include/gfx/hosek_wilkie_sky_model.h:32-34); they are settable here. */
typedef struct hh_sky_params {
    float turbidity;
    float albedo;
    float normalized_sun_y;
    uint32_t cube_size;     /* 512 in the reference (hosek_wilkie_sky_model.cpp:17) */
    /* Extension (not in the reference): > 0 bakes the sky into an
     * equirectangular RGBA16F map of equirect_width x equirect_width/2 texels
     * instead of the cube. Texel (i, j) holds direction
     * (sin t sin p, cos t, -sin t cos p), t = pi (j + 0.5) / H,
     * p = 2 pi ((i + 0.5) / W - 0.5); lookups are bilinear, repeating in i and
     * clamped in j. 0 = the reference cube. */
    uint32_t equirect_width;
} hh_sky_params;

/* CameraNode (include/resource/scene.h:247-282). orientation is a quaternion
 * (x, y, z, w); position = global_position(); fov_y in degrees. */
typedef struct hh_camera {
    float position[3];
    float orientation[4];
    float fov_y_deg;
    float near_plane;
    float far_plane;
    float focal_length;
    float aperture_radius;
} hh_camera;

/* PathIntegrator settings (include/gfx/path_integrator.h:46-51) plus the
 * extensions this build adds. Reference defaults: 7 bounces, 5000 samples,
 * shadow bias 0, untiled. */
typedef struct hh_settings {
    uint32_t max_ray_bounces;
    uint32_t max_samples;
    float shadow_ray_bias;
    uint32_t tiled;
    uint32_t rr_start_depth;   /* 0 = reference behaviour: RR at every bounce */
    uint32_t flags;            /* HH_FLAG_* */
} hh_settings;

#define HH_FLAG_NO_RADIANCE_CLAMP 0x1u   /* diagnostic: skip the min(L, 1) clamp */
#define HH_FLAG_DIRECT_ONLY 0x2u         /* DIRECT_LIGHTING_INTEGRATOR variant */
#define HH_FLAG_MEGAKERNEL 0x4u          /* one kernel per frame instead of the wavefront */
/* Extension (not in the reference, whose area lights are Q2/Q8-quirked):
 * next-event estimation picks uniformly among [all emissive triangles of all
 * instances as one entry, env, directional, point, spot lights], and inside
 * the emissive entry a triangle with probability proportional to its power
 * (luminance(emission) x world area), then a uniform point on it. */
#define HH_FLAG_LIGHT_POWER 0x8u

/* The 192-B PathTraceConsts push-constant block (path_integrator.cpp:11-28). */
typedef struct hh_launch_consts {
    float view_proj_inverse[16];
    float camera_pos[4];
    float up_direction[4];
    float right_direction[4];
    float focal_plane[4];
    int32_t ray_debug_pixel_coord[4];
    uint32_t launch_id_size[4];
    float accumulation;
    uint32_t num_lights;
    uint32_t num_frames;
    uint32_t debug_vis;
    uint32_t max_ray_bounces;
    float shadow_ray_bias;
    float focal_length;
    float aperture_radius;
} hh_launch_consts;

/* A ray and a closest-hit record for the traversal entry points. */
typedef struct hh_ray {
    float origin[3];
    float tmin;
    float direction[3];
    float tmax;
} hh_ray;

typedef struct hh_hit {
    float t;           /* < 0 on miss */
    float u, v;        /* barycentrics of v1, v2 (gl_HitAttribs) */
    uint32_t instance; /* instanceCustomIndex = mesh-node index, 0xFFFFFFFF on miss */
    uint32_t geometry; /* gl_GeometryIndexEXT = submesh */
    uint32_t primitive;/* gl_PrimitiveID within the submesh */
} hh_hit;

#define HH_TRACE_OPAQUE 0x1u            /* gl_RayFlagsOpaqueEXT: skip any-hit */
#define HH_TRACE_TERMINATE_ON_FIRST 0x4u

/* ---------------------------------------------------------------------------
 * Context
 * ------------------------------------------------------------------------- */
const char* hh_last_error(void);
hh_status hh_context_create(int device_ordinal, hh_context* out);
hh_status hh_context_destroy(hh_context ctx);
hh_status hh_sync(hh_context ctx);
/* HIP stream the context launches on (hipStream_t), for interop and events. */
hh_status hh_context_stream(hh_context ctx, void** out_stream);

/* ---------------------------------------------------------------------------
 * Scene (replaces Scene + Mesh + Material + Texture + BatchUploader)
 * ------------------------------------------------------------------------- */
hh_status hh_scene_create(hh_context ctx, hh_scene* out);
hh_status hh_scene_destroy(hh_scene scene);
hh_status hh_scene_add_texture(hh_scene scene, uint32_t width, uint32_t height, uint32_t format,
                               const void* texels, uint32_t* out_id);
hh_status hh_scene_add_material(hh_scene scene, const hh_material_desc* desc, uint32_t* out_id);
/* Mesh::create (mesh.cpp:27-42): vertices/indices/submeshes + the mesh's
 * material list (material ids). Vertex positions must be finite
 * (HH_ERR_INVALID_ARG otherwise). */
hh_status hh_scene_add_mesh(hh_scene scene, const hh_vertex* vertices, uint32_t num_vertices,
                            const uint32_t* indices, uint32_t num_indices,
                            const hh_submesh* submeshes, uint32_t num_submeshes,
                            const uint32_t* material_ids, uint32_t num_materials, uint32_t* out_id);
hh_status hh_scene_add_instance(hh_scene scene, const hh_instance_desc* desc, uint32_t* out_id);
hh_status hh_scene_set_instance_transform(hh_scene scene, uint32_t instance, const float model[16],
                                          const float normal[16]);
hh_status hh_scene_add_light(hh_scene scene, const hh_light_desc* desc, uint32_t* out_id);
/* IBLNode with a cube map: 6 faces of size x size RGBA32F in Vulkan layer
 * order (+X,-X,+Y,-Y,+Z,-Z). */
hh_status hh_scene_set_ibl(hh_scene scene, uint32_t size, const float* rgba32f_faces);
hh_status hh_scene_set_sky(hh_scene scene, const hh_sky_params* params);
/* Packs the GPU records, builds BLAS per mesh + TLAS on the GPU and bakes the
 * sky. Any later edit marks the scene HIERARCHY_UPDATED and needs a commit. */
hh_status hh_scene_commit(hh_scene scene);
hh_status hh_scene_num_lights(hh_scene scene, uint32_t* out);
/* Copies the packed light records (64 B each, scene.cpp:36-42) to host. */
hh_status hh_scene_read_lights(hh_scene scene, float* out_records, uint32_t capacity);
/* Copies the baked environment cube (6 x size x size RGBA32F) to host. */
hh_status hh_scene_read_env(hh_scene scene, float* out_faces, uint32_t* out_size);
/* The equirectangular sky (RGBA16F bits, W x W/2 x 4) when equirect_width > 0;
 * *out_width = 0 when the scene uses the cube. out may be null to query. */
hh_status hh_scene_read_env_equirect(hh_scene scene, uint16_t* out, uint32_t* out_width);
/* Sky coefficients A..I, Z (10 x vec3, as float) used by the bake. */
hh_status hh_scene_sky_coefficients(hh_scene scene, float out[30]);
/* Host-only: the Hosek-Wilkie coefficient fit (A..I, Z as 10 x rgb) for a sun
 * direction, as HosekWilkieSkyModel::update computes it on the CPU
 * (hosek_wilkie_sky_model.cpp:658-701). Needs no device. */
hh_status hh_sky_fit(const hh_sky_params* params, const float sun_dir[3], float out[30]);
/* BVH statistics: node and triangle counts over all BLAS + TLAS. */
hh_status hh_scene_bvh_stats(hh_scene scene, uint64_t out[4]);
/* Traversal limit counters of every ray traced against `scene` since it was
 * committed (or last reset): out[0] pushes dropped from a full per-lane stack,
 * out[1] pushes dropped from a full wave-uniform packet stack, out[2] rays
 * stopped at the bound on node and leaf visits. Each such ray may have missed
 * its closest hit or its occluder; all three stay 0 for trees the builders
 * make. `reset` != 0 clears them after reading. No reference counterpart:
 * the Vulkan driver's traversal has no such limits to report. */
hh_status hh_scene_traversal_faults(hh_scene scene, uint64_t out[3], int reset);

/* ---------------------------------------------------------------------------
 * Scene files: glTF 2.0 (.gltf with external or data-URI buffers, or .glb),
 * the stand-in for ResourceManager's AssetCore loading
 * (src/engine/core/resource_manager.cpp:48-107,196-222,296-490,494-672),
 * whose formats are not part of the reference. The loader hands every
 * resource to a sink (ResourceManager's role); hh_scene_load_gltf binds the
 * sink to a scene, which the caller then commits. Mapping: mesh -> mesh
 * (first use), primitive -> submesh with its material, mesh node -> instance
 * (normal matrix = global transform without scale), baseColorFactor ->
 * albedo_value^(2.2) = factor, metallicRoughnessTexture -> roughness channel
 * 1 / metallic channel 2, factors other than 1 multiplied into a float copy
 * of the texture, alphaMode MASK -> alpha thresholded at alphaCutoff and
 * alpha-tested, BLEND -> transparent, emissiveFactor x
 * KHR_materials_emissive_strength -> emissive, first perspective camera ->
 * info.camera (focal length 8, aperture 0.1: CameraNode's defaults),
 * KHR_lights_punctual -> lights of radius 0. PNG images only. On failure the
 * sink may already hold part of the file. */
typedef struct hh_scene_sink {
    void* user;
    hh_status (*add_texture)(void* user, uint32_t width, uint32_t height, uint32_t format, const void* texels,
                             uint32_t* out_id);
    hh_status (*add_material)(void* user, const hh_material_desc* desc, uint32_t* out_id);
    hh_status (*add_mesh)(void* user, const hh_vertex* vertices, uint32_t num_vertices, const uint32_t* indices,
                          uint32_t num_indices, const hh_submesh* submeshes, uint32_t num_submeshes,
                          const uint32_t* material_ids, uint32_t num_materials, uint32_t* out_id);
    hh_status (*add_instance)(void* user, const hh_instance_desc* desc, uint32_t* out_id);
    hh_status (*add_light)(void* user, const hh_light_desc* desc, uint32_t* out_id);
} hh_scene_sink;
#define HH_GLTF_WARN_TEXCOORD1 0x1u       /* a texture asked for TEXCOORD_1; TEXCOORD_0 was used */
#define HH_GLTF_WARN_NON_TRIANGLES 0x2u   /* point / line primitives were skipped */
typedef struct hh_gltf_info {
    uint32_t num_meshes, num_instances, num_materials, num_textures, num_lights, num_cameras;
    uint32_t has_camera;      /* camera below is valid */
    uint32_t warnings;        /* HH_GLTF_WARN_* */
    uint64_t num_triangles;
    hh_camera camera;         /* the first perspective camera in node order */
} hh_gltf_info;
hh_status hh_gltf_load(const char* path, const hh_scene_sink* sink, hh_gltf_info* info);
hh_status hh_scene_load_gltf(hh_scene scene, const char* path, hh_gltf_info* info);
/* The IBL node's environment from a file (the reference loads a cube map
 * through AssetCore, resource_manager.cpp:261-291,618-637): a Radiance RGBE (.hdr)
 * equirectangular image (value = (m + 0.5) 2^(e - 136); flat or run-length
 * scanlines; "-Y H +X W" or "+Y H +X W"), resampled to 6 faces of
 * size x size RGBA32F in hh_scene_set_ibl's layout. Lat-long texel (i, j) of
 * a W x H image holds direction (sin t sin p, cos t, -sin t cos p),
 * t = pi (j + 0.5) / H, p = 2 pi ((i + 0.5) / W - 0.5); each cube texel
 * averages k x k bilinear lat-long samples over its footprint,
 * k = clamp(ceil(W / (4 size)), 1, 8). out_faces holds 6 size^2 4 floats. */
hh_status hh_ibl_from_hdr(const char* path, uint32_t size, float* out_faces);
hh_status hh_scene_load_ibl_hdr(hh_scene scene, const char* path, uint32_t size);

/* ---------------------------------------------------------------------------
 * Output surface (replaces the RGBA32F ping-pong images)
 * ------------------------------------------------------------------------- */
hh_status hh_surface_create(hh_context ctx, uint32_t width, uint32_t height, hh_surface* out);
hh_status hh_surface_destroy(hh_surface s);
hh_status hh_surface_clear(hh_surface s);  /* vkCmdClearColorImage to (0,0,0,1) */
hh_status hh_surface_read(hh_surface s, float* out_rgba);
hh_status hh_surface_write(hh_surface s, const float* rgba);
hh_status hh_surface_device_ptr(hh_surface s, void** out);
/* Counts non-finite components of the surface; HH_ERR_NAN_DETECTED when any. */
hh_status hh_surface_validate(hh_surface s, uint64_t* nonfinite);
/* First-hit feature planes for the denoiser (no counterpart in the
 * reference). Once enabled, every progressive frame also accumulates, with
 * the radiance's own recurrence, plane 0 = (albedo rgb, 1/(1+t)) and plane 1
 * = (shading normal xyz, M2), M2 being Welford's sum of squared deviations
 * of the clamped sample luminance (Rec. 709) about the running mean. On a
 * miss: albedo = environment radiance clamped to 1, depth 0, normal 0.
 * hh_surface_read_features writes plane 0 then plane 1 (2 x W*H x 4 floats). */
hh_status hh_surface_enable_features(hh_surface s);
hh_status hh_surface_read_features(hh_surface s, float* out);
/* tone_map.frag: exposure -> ACES (op 0) / Reinhard (op 1) / none -> gamma
 * 1/2.2 -> RGBA8 (row 0 = bottom, as stored). */
hh_status hh_tonemap(hh_surface s, float exposure, uint32_t op, uint8_t* out_rgba8);
/* Renderer::save_image_to_disk + copy_and_save_tone_mapped_image
 * (renderer.cpp:637-713, 755-768): hh_tonemap, then an 8-bit RGBA PNG with
 * the top image row first (the reference's tone-map pass flips the stored
 * rows with a negative viewport, renderer.cpp:396-403). */
hh_status hh_surface_save_png(hh_surface s, float exposure, uint32_t op, const char* path);
/* The PNG writer alone (stbi_write_png's role): RGBA8 rows as given, row 0
 * first; flip_rows != 0 writes the last row first. */
hh_status hh_write_png(const char* path, const uint8_t* rgba8, uint32_t width, uint32_t height, uint32_t flip_rows);

/* Output-buffer views (Renderer::OutputBuffer, renderer.h:27-35;
 * render_debug_visualization, renderer.cpp:491-547; debug_visualization.vert
 * and .frag:60-163). Writes, per pixel of `out` (rows in the surface's order:
 * pixel (x, y) sees what the path tracer's pixel (x, y) sees), the closest
 * surface's
 *   ALBEDO    material albedo rgb (as stored, Q10-converted) or texture rgb,
 *   NORMALS   normal * 0.5 + 0.5, the normal being the interpolation of the
 *             per-vertex world normals normalize(mat3(normal_matrix) * N)
 *             (normal-mapped when the material has a map),
 *   ROUGHNESS roughness (not clamped to the path's 0.1 minimum),
 *   METALLIC  metallic,
 *   EMISSIVE  emissive rgb,
 * alpha 1, and (0, 0, 0, 1) where nothing is hit. Matches the raster pass:
 * no alpha test, no culling, view depth clipped to [2fn/(f+n), f] (glm's
 * GL-style projection under Vulkan's 0 <= z <= w clip volume), textures at
 * level 0. FINAL is the path-traced surface itself and is rejected here.
 * Asynchronous on the context stream. */
#define HH_OUTPUT_BUFFER_ALBEDO 0u
#define HH_OUTPUT_BUFFER_NORMALS 1u
#define HH_OUTPUT_BUFFER_ROUGHNESS 2u
#define HH_OUTPUT_BUFFER_METALLIC 3u
#define HH_OUTPUT_BUFFER_EMISSIVE 4u
#define HH_OUTPUT_BUFFER_FINAL 5u
hh_status hh_render_output_buffer(hh_scene scene, const hh_camera* cam, uint32_t output_buffer, hh_surface out);

/* ---------------------------------------------------------------------------
 * Integrator (replaces helios::PathIntegrator)
 * ------------------------------------------------------------------------- */
hh_status hh_integrator_create(hh_context ctx, hh_integrator* out);
hh_status hh_integrator_destroy(hh_integrator it);
hh_status hh_integrator_set_settings(hh_integrator it, const hh_settings* s);
hh_status hh_integrator_get_settings(hh_integrator it, hh_settings* s);
hh_status hh_integrator_restart_bake(hh_integrator it);
/* PathIntegrator::render: one progressive sample (or one tile sample). */
hh_status hh_integrator_render(hh_integrator it, hh_scene scene, const hh_camera* cam, hh_surface out);
/* Renders `frames` consecutive progressive samples in one call (one hipGraph
 * replay per frame). */
hh_status hh_integrator_render_frames(hh_integrator it, hh_scene scene, const hh_camera* cam,
                                      hh_surface out, uint32_t frames);
/* counters: [num_accumulated_samples, num_target_samples, tile_idx, ...] */
hh_status hh_integrator_counters(hh_integrator it, uint32_t out[4]);
/* Path segments traced in the last render call: [extension rays incl.
 * primary, shadow rays, frames]. */
hh_status hh_integrator_ray_stats(hh_integrator it, uint64_t out[3]);
/* Ray-debug path variant (path_trace_debug.rgen; RAY_DEBUG_VIEW branches of
 * path_trace_rgen.glsl:138-148,190-195, path_trace_rchit.glsl:500-514,548-567
 * and path_trace_rmiss.glsl:40-58; PathIntegrator::gather_debug_rays,
 * path_integrator.cpp:88-104; Renderer::add_ray_debug_view /
 * clear_ray_debug_views and the vertex and draw-argument buffers,
 * renderer.cpp:229-249,733-751,1473-1479). A segment buffer holds
 * DebugRayVertex records (common.glsl:54-58) and VkDrawIndirectCommand-style
 * draw arguments {vertex count, 1, 0, 0}; create and clear set them to
 * {0, 1, 0, 0}. */
typedef struct hh_debug_ray_vertex {
    float position[4];
    float color[4];
} hh_debug_ray_vertex;
#define HH_MAX_DEBUG_RAY_DRAW_COUNT 1024u   /* segments; renderer.h:9 */
hh_status hh_debug_rays_create(hh_context ctx, uint32_t max_segments, hh_debug_rays* out);
hh_status hh_debug_rays_destroy(hh_debug_rays r);
hh_status hh_debug_rays_clear(hh_debug_rays r);
/* Traces num_debug_rays paths through window pixel (pixel_x, pixel_y) of a
 * width x height view (window rows top-down: the reference passes
 * (pixel_x, height - pixel_y) to the shader), seeded with the integrator's
 * current num_accumulated_samples and launch id (i, 0), max_ray_bounces and
 * flags from its settings (HH_FLAG_LIGHT_POWER is honoured; the variant is
 * always the full path integrator). Each path draws its colour
 * (3 x (next_float * 0.5 + 0.5)) before the camera ray, skips Russian
 * roulette, and appends one segment (2 vertices) per non-primary ray: origin
 * to hit point, or origin + 10000 * direction on a miss. Appends after what
 * the buffer holds, in no fixed order; the count keeps growing past the
 * capacity, vertices beyond it are dropped. Asynchronous on the context
 * stream. */
hh_status hh_integrator_gather_debug_rays(hh_integrator it, hh_scene scene, const hh_camera* cam, uint32_t width,
                                          uint32_t height, uint32_t pixel_x, uint32_t pixel_y, uint32_t num_debug_rays,
                                          hh_debug_rays out);
/* Copies min(count, capacity, max) vertices and the 4 draw arguments. */
hh_status hh_debug_rays_read(hh_debug_rays r, hh_debug_ray_vertex* out, uint32_t max_vertices, uint32_t draw_args[4]);
hh_status hh_debug_rays_device_ptr(hh_debug_rays r, void** vertices, void** draw_args);

/* PushConstants as launch_rays fills them (path_integrator.cpp:136-159). */
hh_status hh_compute_launch_consts(const hh_camera* cam, uint32_t width, uint32_t height,
                                   uint32_t num_lights, uint32_t num_frames,
                                   const hh_settings* s, hh_launch_consts* out);

/* ---------------------------------------------------------------------------
 * Batched wavefront rendering: many views / scenes of uneven size in one
 * persistent launch sequence per frame.
 * ------------------------------------------------------------------------- */
typedef struct hh_view {
    hh_scene scene;
    hh_surface surface;
    hh_camera camera;
    uint32_t spp;              /* progressive samples per call */
    uint32_t max_ray_bounces;
    float shadow_ray_bias;
    uint32_t first_frame;      /* num_frames of the first sample */
    uint32_t flags;            /* HH_FLAG_NO_RADIANCE_CLAMP / DIRECT_ONLY / LIGHT_POWER */
} hh_view;
/* Renders spp progressive samples of every view, up to frames_in_flight
 * samples per view per wavefront pass (and at most queue_capacity path slots
 * per pass when that is set); each view accumulates in frame order. */
hh_status hh_render_batch(hh_context ctx, const hh_view* views, uint32_t num_views);


/* ---------------------------------------------------------------------------
 * Traversal entry points (closest / any hit with the reference's ray-flag
 * semantics), used by parity tests.
 * ------------------------------------------------------------------------- */
hh_status hh_trace_rays(hh_scene scene, const hh_ray* rays, uint32_t n, uint32_t flags, hh_hit* out_hits);
/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc). hh_trace_rays on caller-owned DEVICE buffers: d_rays holds
 * n hh_ray records (16-byte aligned), d_hits receives n hh_hit records, with
 * the same results as hh_trace_rays. The traversal stacks and raw hits live in
 * a fixed per-GPU scratch array inside the library (about 14.4 MB, held by each
 * GPU it runs on until the process ends); nothing is allocated or freed.
 * n <= HH_TRACE_NOALLOC_MAX_RAYS, larger n returns HH_ERR_LIMIT_EXCEEDED.
 * Asynchronous on the context's stream (hh_sync waits); the scratch is shared
 * by every call on one GPU, so calls that overlap on one GPU race. */
#define HH_TRACE_NOALLOC_MAX_RAYS 32768u
hh_status hh_trace_rays_noalloc(hh_scene scene, const hh_ray* d_rays, uint32_t n, uint32_t flags, hh_hit* d_hits);
/* Primary rays of frame `num_frames` for every pixel, traced with flags 0. */
hh_status hh_primary_hits(hh_scene scene, const hh_camera* cam, uint32_t width, uint32_t height,
                          uint32_t num_frames, hh_hit* out_hits);

/* ---------------------------------------------------------------------------
 * Tuning: queue capacity, BVH width and workgroup shapes.
 * ------------------------------------------------------------------------- */
typedef struct hh_tuning {
    uint32_t bvh_width;       /* 4 or 8 (BLAS/TLAS node width) */
    uint32_t extend_block;    /* threads per block: 64/128/256 */
    uint32_t shade_block;
    uint32_t queue_capacity;  /* rays per wavefront pass; 0 = all paths */
    uint32_t persistent_blocks_per_cu;
    /* Progressive frames traced together by one wavefront pass when
     * hh_integrator_render_frames is asked for several frames. Samples are
     * still accumulated in frame order, so the image is bit-identical to
     * rendering the frames one at a time; deep-bounce queues are F times
     * fuller, which hides the latency tail of incoherent rays. 1..256; the
     * path state takes about 200 B per pixel and frame in flight. */
    uint32_t frames_in_flight;
    /* 1 = quantized 64-B BVH4 nodes (8-bit child boxes, outward rounded);
     * applies to bvh_width 4 and takes effect at the next scene commit. */
    uint32_t bvh_compressed;
    /* Persistent traversal kernels keep one ray per lane in flight and take
     * new rays from the queue for a wave once this many of its 64 lanes have
     * finished (1..64; 64 = only when the whole wave is idle). */
    uint32_t refill_lanes;
    /* Step shape of the traversal loop: 0 = while-while (inner nodes until
     * the lane holds a leaf, then the leaf), 1 = if-if (one node or leaf per
     * step). Results are identical; only the schedule differs. */
    uint32_t traversal;
    /* BLAS / TLAS builder, used at the next scene commit: 3 = binned SAH on
     * the GPU, every mesh's tree in one build (fast trace, the role of
     * PREFER_FAST_TRACE in mesh.cpp:109; the default); 1 = the same binned
     * SAH on the host (the same splits, hence the same tree); 0 = LBVH on the
     * GPU (Morton order); 2 = PLOC clustering on the GPU (Meister and Bittner
     * 2018; radius 16). Hits are identical either way (ties resolve by
     * instance, geometry, primitive). */
    uint32_t bvh_builder;
    /* Paths still alive after this many bounces traced as wavefront launches
     * finish in one continuation launch, in which a lane whose path ends takes
     * the next queued path (deep bounces hold few rays, and a launch per bounce
     * pays its latency tail each time). 0 = every bounce as wavefront launches,
     * 1..64 otherwise; HH_FLAG_MEGAKERNEL (the single-frame form) always
     * continues after the camera rays' bounce, i.e. 1. Results are identical. */
    uint32_t continuation_depth;
} hh_tuning;
hh_status hh_context_set_tuning(hh_context ctx, const hh_tuning* t);
hh_status hh_context_get_tuning(hh_context ctx, hh_tuning* t);
/* Autotuner (the reference leaves all of this to its driver): times
 * `frames` progressive frames of the views, rendered into scratch surfaces
 * (the views' own surfaces are untouched), under candidate tunings and keeps
 * the fastest on the context: coordinate descent, twice over, on
 * frames_in_flight {1,2,4,8,16}, extend_block {64,128,256},
 * persistent_blocks_per_cu {4,8,16,32}, refill_lanes {16,32,48,64} and
 * traversal {0,1}; with HH_AUTOTUNE_BVH_WIDTH it also re-commits the views'
 * scenes with the other BVH width and keeps the faster (a re-commit restarts
 * the progressive accumulation of integrators on those scenes). Tuning never
 * changes an image (tested bit-identical). Returns the chosen tuning and its
 * milliseconds per frame. */
#define HH_AUTOTUNE_BVH_WIDTH 0x1u
hh_status hh_context_autotune(hh_context ctx, const hh_view* views, uint32_t num_views, uint32_t frames, uint32_t flags,
                              hh_tuning* out, double* out_ms_per_frame);

/* ---------------------------------------------------------------------------
 * Measurement (replaces the reference's scoped GPU timestamps,
 * src/engine/utility/profiler.cpp:77-146). flags: bit 0 = HIP events around
 * every kernel launch (frames are then launched directly, not from a graph);
 * bit 1 = counting traversal variant (node / triangle / instance fetches).
 * Kernel kinds: 0 generate, 1 extend, 2 shade, 3 shadow, 4 accumulate,
 * 5 megakernel. counts: extension-ray nodes, triangles, instances, then the
 * same for shadow rays.
 * ------------------------------------------------------------------------- */
hh_status hh_context_set_profiling(hh_context ctx, uint32_t flags);
hh_status hh_context_read_profile(hh_context ctx, double ms[6], uint64_t launches[6], uint64_t counts[6]);
hh_status hh_context_reset_profile(hh_context ctx);
/* With timing on: extension and shadow rays traced by the launches of each
 * kernel kind since the last reset (as of the last hh_context_read_profile),
 * so per-launch work matches the timed launches exactly. */
hh_status hh_context_profile_rays(hh_context ctx, uint64_t extension[6], uint64_t shadow[6]);
/* Streams `bytes` through a plain copy kernel `iters` times and returns the
 * achieved read+write GB/s: the calibration point for HBM byte counters. */
hh_status hh_bandwidth_probe(hh_context ctx, uint64_t bytes, uint32_t iters, double* out_gbps);

/* ---------------------------------------------------------------------------
 * Learned denoiser (no counterpart in the reference, which only accumulates;
 * plain progressive accumulation stays the fallback and the reference of its
 * error). A U-Net with levels of 32/48/64/96 channels: per level two 3x3
 * convolutions + ReLU, 2x2 average pooling down, nearest upsampling and skip
 * concatenation up, and a 1x1 head whose output is added to the input's
 * log(1 + radiance). BF16 weights and activations on MFMA, FP32
 * accumulation, FP32 master weights trained with Adam.
 *
 * Inputs per pixel (the network's 10 channels): log(1 + radiance),
 * albedo, octahedral(normal), 1/(1+t), and the standard deviation of the mean
 * luminance sqrt(M2 / (n (n-1))) from the surface's feature planes.
 * Raw feature records (training and tests) hold 12 floats per pixel:
 * radiance rgb, albedo rgb, 1/(1+t), normal xyz, M2, n (samples held).
 * Images of any size are padded with zeros to multiples of 8.
 *
 * Parameters are one FP32 vector: per layer W (Cout, 3, 3, Cin) then b
 * (Cout), for the 14 convolutions in order (enc0a, enc0b, enc1a, enc1b,
 * enc2a, enc2b, mida, midb, dec2a, dec2b, dec1a, dec1b, dec0a, dec0b), then
 * the head W (3, 32) and b (3): 480,675 floats.
 * ------------------------------------------------------------------------- */
hh_status hh_denoiser_create(hh_context ctx, uint64_t seed, hh_denoiser* out);  /* He-normal init */
hh_status hh_denoiser_destroy(hh_denoiser dn);
hh_status hh_denoiser_num_params(uint64_t* out);
hh_status hh_denoiser_set_params(hh_denoiser dn, const float* params, uint64_t n);  /* resets Adam */
hh_status hh_denoiser_get_params(hh_denoiser dn, float* params, uint64_t n);
/* Gradient of the last hh_denoiser_loss_grad / train_step call. */
hh_status hh_denoiser_get_grads(hh_denoiser dn, float* grads, uint64_t n);
/* Denoises a surface with feature planes after `num_frames` progressive
 * frames (the running mean then holds max(num_frames - 1, 1) samples) into
 * `out` (same size, RGBA32F, alpha 1). */
hh_status hh_denoise(hh_denoiser dn, hh_surface in, uint32_t num_frames, hh_surface out);
/* Host-array entry points: raw records (n x h x w x 12) in, the network
 * output in log(1 + radiance) space (n x h x w x 3) out. */
hh_status hh_denoiser_forward(hh_denoiser dn, const float* feats, uint32_t n, uint32_t h, uint32_t w, float* out_log);
/* Training loss, mean over pixels and channels (target: n x h x w x 3):
 *   HH_DENOISER_LOSS_LOG_MSE (default)  (out - log(1 + target))^2
 *   HH_DENOISER_LOSS_RELMSE             (y - t)^2 / (t^2 + 0.01), y = exp(out) - 1
 * hh_denoiser_loss_grad returns the loss and keeps its gradient (read with
 * hh_denoiser_get_grads). */
#define HH_DENOISER_LOSS_LOG_MSE 0u
#define HH_DENOISER_LOSS_RELMSE 1u
hh_status hh_denoiser_set_loss(hh_denoiser dn, uint32_t kind);
hh_status hh_denoiser_loss_grad(hh_denoiser dn, const float* feats, const float* target, uint32_t n, uint32_t h, uint32_t w,
                                float* loss);
/* loss_grad + one Adam step (beta 0.9 / 0.999, eps 1e-8) at learning rate lr. */
hh_status hh_denoiser_train_step(hh_denoiser dn, const float* feats, const float* target, uint32_t n, uint32_t h, uint32_t w,
                                 float lr, float* loss);

#ifdef __cplusplus
}
#endif
#endif /* HELIOS_HIP_H */
