/*
 * nemo_flash.h -- FLASH-format HDF5 checkpoint writer (host side).
 *
 * Replaces the reference's hook_checkpoint_{100,200,300,500}
 * (checkpoint/flash/checkpoint_mod.f90:11-364) with the datasets of
 * setups/euler/3d/_common_/setup_checkpoint_mod.f90:9-60.  The caller builds
 * the block tree (mesh_alloc_tree, mesh/p4est/mesh_utils_mod.f90:443-586:
 * the root, then the parent blocks level by level in Morton order, then
 * every leaf in Morton order) and passes the leaf states; nemo_amd/flash.py
 * does this for the adaptive forest and the uniform meshes.
 */
#ifndef NEMO_FLASH_H
#define NEMO_FLASH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NEMO_FLASH_OK 0
#define NEMO_FLASH_ERR 1

typedef struct {
    int32_t nxb;                  /* nodes per block and direction (N_NODES) */
    int32_t nstep;                /* rt%timesteps */
    int32_t file_number;          /* rt%chkptcount */
    int32_t lrefine_min;          /* mesh%minlevel + 1 */
    int32_t lrefine_max;          /* mesh%maxlevel + 1 */
    int32_t file_format_version;  /* 9 (checkpoint_types_mod.f90:61) */
    int8_t boundary[6];           /* box sides x-, x+, y-, y+, z-, z+: 0 periodic, 1 outflow,
                                     2 reflecting (the xl_boundary_type .. zr_boundary_type strings) */
    int16_t pad;
    double time, dt, next_checkpoint_time;
    double xmin, xmax, ymin, ymax, zmin, zmax;
    double cfl, gamma, tinitial, tmax, checkpoint_interval_time;
} nemo_flash_meta;

typedef struct {
    int64_t numblocks;             /* tree blocks: parents and leaves */
    int64_t nleaves;               /* the last nleaves blocks are the leaves */
    const double *coordinates;     /* [numblocks][3] block centres */
    const double *block_size;      /* [numblocks][3] */
    const double *bounding_box;    /* [numblocks][3][2] lower, upper */
    const int32_t *gid;            /* [numblocks][15]: 6 faces, parent, 8 children (1-based, -1 none) */
    const int32_t *node_type;      /* [numblocks]: 1 leaf, 2 parent */
    const int32_t *refine_level;   /* [numblocks]: level + 1 */
    const int32_t *which_child;    /* [numblocks]: 1..8, -1 for the root */
    const double *dtmin;           /* optional [nleaves]: per-leaf CFL time step (/dtmin) */
} nemo_flash_tree;

/* Write <path> (through <path>.partial and a rename).  unknown_names: nunk
 * names of 4 characters each ("dens", "momx", ...); leaf_data:
 * [nunk][nleaves][nxb^3], node i + nxb*j + nxb^2*k of a leaf; leaf_blend:
 * optional [nleaves][nxb^3] blending factor (/blnd).  Returns
 * NEMO_FLASH_OK or NEMO_FLASH_ERR (message: nemo_flash_error_message). */
int nemo_flash_write(const char *path, const nemo_flash_meta *meta, const nemo_flash_tree *tree, int32_t nunk,
                     const char *unknown_names, const double *leaf_data, const double *leaf_blend);
int nemo_flash_error_message(char *buf, int32_t len);

#ifdef __cplusplus
}
#endif
#endif
