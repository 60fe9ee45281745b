/*
 * ha_mesh.h -- geodesic searchlights on a triangle mesh.
 *
 * No C counterpart in the reference: its real-data benchmark takes the
 * searchlights from neuroboros (benchmark_neuroboros.py:153, nb.sls(lr,
 * radius, return_dists=True)) and passes them to the searchlight entry points
 * as flat arrays (hyperalignment_c.py:161-169). This builds those arrays from
 * any triangle mesh, on the host.
 */
#ifndef HA_MESH_H
#define HA_MESH_H

#include "ha_common.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Searchlights of radius `radius` (mesh units, mm for cortical surfaces) around
 * each centre (centers, n_centers; NULL = every vertex): the vertices whose
 * distance along the mesh edges (Dijkstra; edge length = Euclidean distance of
 * its end points) is below radius, the centre included at 0, sorted by vertex
 * id. coords: nv x 3 (row-major), faces: nf x 3 vertex ids. Outputs, in the flat
 * layout of the searchlight entry points, are malloc'ed here and released with
 * ha_mesh_free: sl_indices (sl_offsets[n] entries), sl_offsets (n + 1),
 * sl_dists (distances, aligned with sl_indices). */
int ha_mesh_searchlights(const double *coords, int32_t nv, const int32_t *faces, int32_t nf, double radius,
                         const int32_t *centers, int32_t n_centers, int32_t **sl_indices, int32_t **sl_offsets,
                         double **sl_dists);
/* The same lists with geodesic distances across the faces instead of along the
 * edges: fast marching with the planar unfolding update (a face whose two
 * other vertices are accepted gives the third the distance from the virtual
 * point source at their two distances, when the straight ray crosses their
 * edge; every edge also gives its Dijkstra update). Exact on flat meshes whose
 * geodesics cross the faces; on a sphere within a small fraction of the
 * great-circle distance (tests/test_mesh.py). */
int ha_mesh_searchlights_geodesic(const double *coords, int32_t nv, const int32_t *faces, int32_t nf,
                                  double radius, const int32_t *centers, int32_t n_centers, int32_t **sl_indices,
                                  int32_t **sl_offsets, double **sl_dists);
void ha_mesh_free(void *p);

#ifdef __cplusplus
}
#endif

#endif /* HA_MESH_H */
