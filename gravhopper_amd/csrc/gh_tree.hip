// Barnes-Hut octree on the device (placeholder until the build and walk land).
#include "gh_internal.h"

namespace gh {

struct TreeBufs {};

int tree_dev(gh_ctx* c, const double*, const double*, int64_t, const double*, int64_t, double,
             double, int, double*, hipStream_t)
{
    return fail(c, GH_EARG, "tree force not available in this build");
}

void tree_free(gh_ctx* c)
{
    delete c->tree;
    c->tree = nullptr;
}

}  // namespace gh

extern "C" int gh_tree_dump(gh_ctx* c, const double*, const double*, int64_t, double,
                            gh_node_rec*, int64_t, int64_t*)
{
    return gh::fail(c, GH_EARG, "tree dump not available in this build");
}
