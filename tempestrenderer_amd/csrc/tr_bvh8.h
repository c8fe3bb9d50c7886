/*
 * tr_bvh8.h - host builder of the 8-wide BVH.
 */
#ifndef TR_BVH8_H
#define TR_BVH8_H

#include <string>
#include <vector>
#include "tr_scene.h"

namespace trk {

struct Bvh8Stats {
    uint32_t depth = 0;
    uint32_t nodes = 0;
};

/* Builds Node8 nodes over `prims`; `prims_out` receives the primitives in
 * leaf order (the order the nodes' prim_base/meta refer to). */
bool build_bvh8(const std::vector<PrimRec>& prims, std::vector<Node8>& nodes_out, std::vector<PrimRec>& prims_out,
                Bvh8Stats* stats, std::string* err);

}  // namespace trk

#endif
