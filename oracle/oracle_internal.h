/*
 * oracle_internal.h — shared internals of the CPU oracle (TEST INFRASTRUCTURE
 * ONLY; see oracle.h).
 */
#ifndef WL_ORACLE_INTERNAL_H
#define WL_ORACLE_INTERNAL_H

#include "oracle.h"
#include <math.h>
#include <stddef.h>

typedef struct
{
    float lo[3], hi[3];
    uint32_t first; /* leaf: first index into prims; inner: left child id */
    uint16_t count; /* leaf: number of prims; inner: 0                     */
    uint16_t axis;  /* inner: split axis (child order hint)                */
} obvh_node;

struct oracle_bvh
{
    obvh_node *nodes;
    uint32_t nnodes;
    uint32_t *prims;
    float *tris; /* copy, 9 floats per triangle (global prim order) */
    uint32_t ntris;
    uint32_t depth;
};

struct oracle_pt
{
    uint32_t num_meshes;
    float *vertices;
    float *normals;
    uint32_t *indices;
    uint32_t *mesh_vertex_offset;
    uint32_t *mesh_tri_offset;
    uint32_t *mesh_tri_count;
    uint32_t *mesh_has_normal;
    uint32_t *mesh_light;
    oracle_material *mesh_material;
    uint32_t num_lights;
    uint32_t *light_mesh;
    float *light_L;
    uint32_t *light_two_sided;
    float cam_pos[3], cam_lookat[3], cam_up[3], cam_right[3];
    float bg[3];
    uint32_t num_textures;
    float *texcoords;
    uint32_t *mesh_texture;
    uint32_t *texture_offset;
    uint32_t *texture_size;
    uint8_t *texels;
    oracle_bvh *bvh;
};

/* Run fn(ctx, i) for i in [0, n) on nthreads threads (dynamic chunks). */
typedef void (*oracle_task_fn)(void *ctx, uint32_t begin, uint32_t end);
void oracle_parallel_for(uint32_t n, uint32_t chunk, int nthreads,
                         oracle_task_fn fn, void *ctx);

#endif
