/*
 * atrbvh_scenes.h -- seeded synthetic scene and schedule utilities that ship
 * with the library.  They have no counterpart in the reference (whose scenes
 * are OBJ files loaded by its Windows framework); they produce the benchmark
 * and test workloads.
 *
 * Random numbers: SplitMix64 seeding a PCG32 (XSH-RR) stream; uniform floats
 * are (u32 >> 8) * 2^-24; normal deviates use Box-Muller in double and are
 * rounded to fp32 once.  Vertex arrays are float4[3n] with w = 0.
 */
#ifndef ATRBVH_SCENES_H
#define ATRBVH_SCENES_H

#include <stdint.h>

#include "atrbvh.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Triangle soup: centroids uniform in [0,1)^3, each vertex = centroid +
 * N(0, sigma) per axis. */
void atr_gen_soup(uint64_t seed, uint32_t n, float sigma, float* vertices);

/* Gaussian-mixture soup: `clusters` centres uniform in [0,1)^3 with per-cluster
 * sigma log-uniform in [sigma_lo, sigma_hi]; each triangle picks a cluster
 * uniformly, its centroid ~ N(centre, sigma^2 I), and is an equilateral
 * triangle of side L ~ LogNormal(ln edge_median, edge_sigma) under a uniform
 * random rotation (normalised Gaussian quaternion). */
void atr_gen_gmm(uint64_t seed, uint32_t n, uint32_t clusters, float sigma_lo, float sigma_hi,
                 float edge_median, float edge_sigma, float* vertices);

/* Mesh sizes of a batch: N_m = round(exp(U[ln lo, ln hi])) clamped to
 * [lo, hi]; writes offsets[0..meshes] (prefix sums, offsets[0] = 0). */
void atr_gen_batch_sizes(uint64_t seed, uint32_t meshes, uint32_t lo, uint32_t hi,
                         uint32_t* offsets);

/* Batch of soups: mesh m is atr_gen_soup(seed_base + m, N_m, sigma) written at
 * triangle offsets[m] (host threads, deterministic). */
void atr_gen_soup_batch(uint64_t seed_base, uint32_t meshes, const uint32_t* offsets, float sigma,
                        float* vertices);

/* Wavefront OBJ scenes, read as the reference framework's importer reads
 * them (gpu-ray-traversal/src/framework/io/MeshWavefrontIO.cpp:330-545):
 * "v" positions (exactly three numbers, parsed with the framework's fp32
 * accumulation, String.cpp:440-491), "f" faces ("p", "p/t", "p/t/n",
 * "p//n"; negative indices count from the end; out-of-range positions read
 * as (0,0,0)) triangulated as fans, faces grouped by the materials of
 * "mtllib" files and listed group by group in order of first use, as the
 * reference's Scene lists them (rt/Scene.cpp:55-77).  Host only. */
typedef struct atr_mesh atr_mesh;
atr_status atr_mesh_load_obj(const char* path, atr_mesh** out);
atr_status atr_mesh_info(const atr_mesh* mesh, uint32_t* num_triangles, uint32_t* num_positions);
/* The triangle soup, float4[3 * num_triangles] with w = 0. */
atr_status atr_mesh_soup(const atr_mesh* mesh, float* vertices);
atr_status atr_mesh_destroy(atr_mesh* mesh);
/* SceneConverter::ConvertToCudaSceneDeviceMemory
 * (gpu-ray-traversal/src/rt/SceneConverter.cpp:19-71): the soup and its box
 * (running maximum starting at FLT_MIN) computed on the host, then
 * atr_scene_create. */
atr_status atr_scene_create_from_mesh(atr_context* ctx, const atr_mesh* mesh, atr_scene** out);

/* TRBVH lane schedule (the library's own generator): writes rounds*warp ints
 * into out when cap allows and returns the number of rounds, or -1. */
int atr_trbvh_schedule(int treelet, int warp, int32_t* out, int cap);

#ifdef __cplusplus
}
#endif

#endif
