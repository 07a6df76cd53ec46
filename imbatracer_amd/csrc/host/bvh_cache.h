// Acceleration-structure cache files (loaders/load_bvh.cpp container), host side.
#pragma once

#include <vector>

#include "imba_hip.h"

namespace imba {

int cache_has_block(const char* path, int type);
// Stores nodes / Vec4 records of one tree, rebased by -node_offset, -tris_offset, -tri_id_offset.
int cache_store(const char* path, int type, const void* nodes, int node_count, const ImbaVec4* tris, int vec4_count,
                int node_offset, int tris_offset, int tri_id_offset);
// Loads the block of `type`, rebased by +node_offset, +tris_offset, +tri_id_offset.
int cache_load(const char* path, int type, int node_offset, int tris_offset, int tri_id_offset,
               std::vector<char>& nodes, std::vector<ImbaVec4>& tris);

}  // namespace imba
