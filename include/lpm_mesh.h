/*
 * lpm_mesh.h -- C ABI of the host mesh builder in liblpm (namespace Lpm).
 *
 * Materializes the particle sets of the spherical sum: the vertices and the
 * leaf-face centroids and areas of a uniformly refined cubed-sphere or
 * icosahedral mesh, in the reference's index order
 * (MeshedParticles(seed, level, ...) src/LpmMeshedParticles.cpp:12-21,
 *  PolyMesh2d src/LpmPolyMesh2d.cpp:20-93).  For FFI callers that cannot use
 * the C++ classes.
 */
#ifndef LPM_MESH_H
#define LPM_MESH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LPM_SEED_ICOS_TRI 0  /* IcosTriSphereSeed, meshSeeds/icosTriSphereSeed.dat */
#define LPM_SEED_CUBED 1     /* CubedSphereSeed, meshSeeds/cubedSphereSeed.dat */

typedef struct lpm_mesh lpm_mesh;

/* returns 0 or a negative lpm_status (-1 invalid argument, -4 allocation) */
int lpm_mesh_create(lpm_mesh** out, int seed, int level, double radius);
void lpm_mesh_destroy(lpm_mesh* m);
/* counts: vertices, all faces (parents included), leaf faces */
int lpm_mesh_sizes(const lpm_mesh* m, int64_t* n_vertices, int64_t* n_faces, int64_t* n_leaf_faces);
int lpm_mesh_vertices(const lpm_mesh* m, double* x, double* y, double* z);
/* leaf faces in ascending face index: centroid, area, face index */
int lpm_mesh_leaf_faces(const lpm_mesh* m, double* cx, double* cy, double* cz, double* area, int64_t* face_index);
/* all faces: area (0 for divided faces), divided flag, centroid */
int lpm_mesh_faces(const lpm_mesh* m, double* area, int8_t* divided, double* cx, double* cy, double* cz);
/* sqrt(max |v|^2) over vertices (src/LpmCoords.cpp:136-143) */
double lpm_mesh_max_radius(const lpm_mesh* m);

#ifdef __cplusplus
}
#endif

#endif /* LPM_MESH_H */
