/*
 * vgi_workloads.h -- seeded synthetic workloads for the tests and the
 * benchmark (libvgi_workloads.so). Not part of the drop-in boundary: a host
 * of the library streams its own pages; these generators only make the
 * benchmark's scenes, optics and edit streams reproducible.
 */
#ifndef VGI_WORKLOADS_H
#define VGI_WORKLOADS_H
#include "vgi.h"
#ifdef __cplusplus
extern "C" {
#endif

/* Dense u16 grids of `dims` voxels, dims multiples of 32. */
/* Workloads: 1 hollow box 64^3, 2 noise terrain 256^3, 4 edit-stream terrain
 * 512^3, 5 worm cave 512^3, 6 furnace (reference research.rs:71-115). */
int vgi_scene_generate(int workload, uint32_t seed, uint32_t edge, uint16_t *voxels, int32_t dims[3]);
/* Materials and lights of a workload (256x3 diffuse, 256x3 emission, lights). */
int vgi_scene_optics(int workload, uint32_t seed, float *diffuse_rgb, float *emission_rgb, vgi_light *lights,
                     uint32_t *light_count, float environment[3]);
/* Exposed faces of a dense grid (solid voxel, air neighbour inside the grid),
 * optionally limited to the box interior; returns the count. */
int64_t vgi_scene_faces(const uint16_t *voxels, const int32_t dims[3], int interior_only, vgi_face_task *out,
                        int64_t capacity);
/* One frame of the edit stream over an edge^3 box: `count` seeded edits drawn
 * from PCG(seed, frame, i), each a voxel set (material 1-4), a voxel clear,
 * or a sphere removal of radius U[1,5] (every voxel whose centre lies within
 * the radius), expanded to voxel edits in order; returns the voxel count
 * (copies at most `capacity`: xyz 3 x i32 and value u16 per voxel). */
int64_t vgi_scene_edits(uint32_t seed, uint32_t frame, uint32_t count, uint32_t edge, int32_t *xyz, uint16_t *values,
                        int64_t capacity);

/* A depth/normal prepass of a dense grid seen from a look-at camera (up
 * +y, vertical field of view fov_y radians, reverse-Z infinite perspective
 * with near plane 0.1: depth = 0.1 / view distance along the axis, 0 where
 * the ray leaves the grid): the first solid voxel along each pixel centre's
 * ray (3D DDA), the hit point's depth and the hit face's normal. Writes
 * world_from_clip (16 floats, column-major, vgi_view_camera's) and per pixel
 * row-major depth and normal xyz. Returns the pixels that hit. */
int64_t vgi_scene_prepass(const uint16_t *voxels, const int32_t dims[3], const float eye[3], const float target[3], float fov_y,
                          uint32_t width, uint32_t height, float *world_from_clip, float *depth, float *normals);

#ifdef __cplusplus
}
#endif
#endif
