// extern "C" mesh builder over Lpm::PolyMesh2d (include/lpm_mesh.h).
#include "LpmMesh.hpp"
#include "lpm_mesh.h"

#include <new>

struct lpm_mesh {
    std::unique_ptr<Lpm::PolyMesh2d> mesh;
};

extern "C" {

int lpm_mesh_create(lpm_mesh** out, int seed, int level, double radius) {
    if (!out) return -2;
    *out = nullptr;
    if (level < 0 || level > 12 || !(radius > 0.0)) return -1;
    try {
        std::unique_ptr<Lpm::MeshSeed> s;
        if (seed == LPM_SEED_ICOS_TRI) s.reset(new Lpm::IcosTriSphereSeed());
        else if (seed == LPM_SEED_CUBED) s.reset(new Lpm::CubedSphereSeed());
        else return -1;
        auto* m = new lpm_mesh();
        m->mesh.reset(new Lpm::PolyMesh2d(*s, level, radius));
        *out = m;
        return 0;
    } catch (const std::bad_alloc&) {
        return -4;
    } catch (...) {
        return -1;
    }
}

void lpm_mesh_destroy(lpm_mesh* m) { delete m; }

int lpm_mesh_sizes(const lpm_mesh* m, int64_t* nv, int64_t* nf, int64_t* nl) {
    if (!m) return -2;
    if (nv) *nv = m->mesh->nVertices();
    if (nf) *nf = m->mesh->nFaces();
    if (nl) *nl = m->mesh->nLeafFaces();
    return 0;
}

int lpm_mesh_vertices(const lpm_mesh* m, double* x, double* y, double* z) {
    if (!m || !x || !y || !z) return -2;
    const auto& p = *m->mesh;
    for (Lpm::index_type i = 0; i < p.nVertices(); ++i) { x[i] = p.vx[i]; y[i] = p.vy[i]; z[i] = p.vz[i]; }
    return 0;
}

int lpm_mesh_leaf_faces(const lpm_mesh* m, double* cx, double* cy, double* cz, double* area, int64_t* idx) {
    if (!m || !cx || !cy || !cz || !area) return -2;
    const auto& p = *m->mesh;
    int64_t k = 0;
    for (Lpm::index_type i = 0; i < p.nFaces(); ++i) {
        if (p.faceIsDivided(i)) continue;
        const Lpm::XyzVector c = p.faceCentroid(i);
        cx[k] = c.x; cy[k] = c.y; cz[k] = c.z;
        area[k] = p.faceArea(i);
        if (idx) idx[k] = i;
        ++k;
    }
    return 0;
}

int lpm_mesh_faces(const lpm_mesh* m, double* area, int8_t* divided, double* cx, double* cy, double* cz) {
    if (!m) return -2;
    const auto& p = *m->mesh;
    for (Lpm::index_type i = 0; i < p.nFaces(); ++i) {
        if (area) area[i] = p.faceArea(i);
        if (divided) divided[i] = p.faceIsDivided(i) ? 1 : 0;
        if (cx && cy && cz) {
            const Lpm::XyzVector c = p.faceCentroid(i);
            cx[i] = c.x; cy[i] = c.y; cz[i] = c.z;
        }
    }
    return 0;
}

double lpm_mesh_max_radius(const lpm_mesh* m) { return m ? m->mesh->maxR() : 0.0; }

}  // extern "C"
