// ff_mesh.hip -- incremental marching cubes over changed bricks (placeholder
// until the cell mesher lands: reports no geometry).
#include "ff_internal.hpp"

namespace ff {
struct MeshState {
    int dummy;
};
int mesh_init(ff_ctx *c) { c->mesh = new MeshState(); return FF_OK; }
void mesh_free(ff_ctx *c) { delete c->mesh; c->mesh = nullptr; }
void mesh_on_initial_volume(ff_ctx *) {}
void mesh_on_grow(ff_ctx *) {}
void mesh_after_frame(ff_ctx *) {}
}  // namespace ff

extern "C" {
int ff_update_meshes(ff_ctx *c) { return c ? 0 : FF_E_ARG; }
int ff_mesh_sizes(ff_ctx *c, size_t *nv, size_t *nf)
{
    if (!c) return FF_E_ARG;
    if (nv) *nv = 0;
    if (nf) *nf = 0;
    return FF_OK;
}
int ff_copy_mesh(ff_ctx *c, float *, uint8_t *, uint32_t *) { return c ? FF_OK : FF_E_ARG; }
int ff_cell_count(ff_ctx *c, size_t *n)
{
    if (!c || !n) return FF_E_ARG;
    *n = 0;
    return FF_OK;
}
int ff_copy_cells(ff_ctx *c, int32_t *, uint32_t *, uint32_t *) { return c ? FF_OK : FF_E_ARG; }
int ff_set_incremental_meshing(ff_ctx *c, int on)
{
    if (!c) return FF_E_ARG;
    c->incremental = on != 0;
    return FF_OK;
}
}
