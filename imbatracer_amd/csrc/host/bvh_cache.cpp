// Acceleration-structure cache files: the container of loaders/load_bvh.cpp.
//
//   file  = u32 magic 0x313F1A57, then blocks
//   block = u64 size (bytes after this field), u32 type (1: BVH2 GPU layout, 2: BVH4 CPU
//           layout), u32 node_count, u32 prim_count (Vec4 records), the node records, the
//           Vec4 records
//
// One file holds at most one block per type (store refuses to add a second; load_bvh.cpp:
// 170-177).  A block keeps one mesh's tree with links relative to its own first node and first
// Vec4 and triangle ids relative to the mesh's first global id; load rebases them to where the
// caller puts them (load_bvh.cpp:66-95, 121-139) and store does the inverse (:187-226,
// 251-288).  The node records are this library's 64-byte BVH2 and 128-byte BVH4 nodes.  Two
// rules differ from the reference loader: the dummy child 0x76543210 of a single-leaf tree is
// kept as is (the reference shifts it along with node links), and a packet sentinel is a Vec4
// with all four lanes 0x80000000 (the reference looks at lane x only).
#include <cstdio>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "bvh_cache.h"
#include "status.h"

namespace imba {

namespace {

constexpr uint32_t kMagic = 0x313F1A57u;
constexpr uint32_t kSentinel = 0x80000000u;

uint32_t bits(float f) {
    uint32_t b;
    std::memcpy(&b, &f, 4);
    return b;
}
float from_bits(uint32_t b) {
    float f;
    std::memcpy(&f, &b, 4);
    return f;
}

size_t node_bytes(int type) { return type == IMBA_CACHE_BVH2 ? sizeof(ImbaNode) : sizeof(ImbaNode4); }

bool is_sentinel(const ImbaVec4& v) {
    return bits(v.x) == kSentinel && bits(v.y) == kSentinel && bits(v.z) == kSentinel && bits(v.w) == kSentinel;
}

// Positions the stream after the type field of the first block of `type`; false if absent.
bool locate(std::ifstream& in, uint32_t type, uint64_t* size) {
    for (;;) {
        uint64_t sz = 0;
        uint32_t t = 0;
        if (!in.read((char*)&sz, 8) || !in.read((char*)&t, 4)) return false;
        if (sz < 4) return false;
        if (t == type) {
            *size = sz;
            return true;
        }
        in.seekg((std::streamoff)(sz - 4), std::ios::cur);
        if (!in) return false;
    }
}

bool has_magic(std::ifstream& in) {
    uint32_t m = 0;
    return in.read((char*)&m, 4) && m == kMagic;
}

// Rebases links and ids by +sign * offsets.
void rebase_bvh2(ImbaNode* n, int count, int node_off, int tris_off) {
    for (int i = 0; i < count; i++) {
        for (int* c : {&n[i].left, &n[i].right}) {
            if (*c == IMBA_DUMMY_CHILD && c == &n[i].right) continue;
            *c = *c < 0 ? ~(~*c + tris_off) : *c + node_off;
        }
    }
}

void rebase_bvh4(ImbaNode4* n, int count, int node_off, int tris_off) {
    for (int i = 0; i < count; i++)
        for (int& c : n[i].children)
            if (c > 0) c += node_off;
            else if (c < 0) c = ~(~c + tris_off);
}

int rebase_ids_bvh2(ImbaVec4* t, int count, int id_off) {
    if (count % 3) return fail(IMBA_ERR_BAD_BVH, "BVH2 triangle records are not a multiple of 3 Vec4");
    for (int i = 0; i < count; i += 3) t[i + 1].w = from_bits(bits(t[i + 1].w) + (uint32_t)id_off);
    return IMBA_OK;
}

int rebase_ids_bvh4(ImbaVec4* t, int count, int id_off) {
    int i = 0;
    while (i < count) {
        if (is_sentinel(t[i])) {
            i++;
            continue;
        }
        if (i + 13 > count) return fail(IMBA_ERR_BAD_BVH, "truncated triangle packet at Vec4 %d", i);
        float* ids = &t[i + 12].x;
        for (int k = 0; k < 4; k++)
            if (bits(ids[k]) != kSentinel) ids[k] = from_bits(bits(ids[k]) + (uint32_t)id_off);
        i += 13;
    }
    return IMBA_OK;
}

}  // namespace

int cache_has_block(const char* path, int type) {
    std::ifstream in(path, std::ios::binary);
    uint64_t sz;
    return in && has_magic(in) && locate(in, (uint32_t)type, &sz);
}

int cache_store(const char* path, int type, const void* nodes, int node_count, const ImbaVec4* tris, int vec4_count,
                int node_offset, int tris_offset, int tri_id_offset) {
    IMBA_CHECK_ARG(path && (nodes || node_count == 0) && (tris || vec4_count == 0), "NULL argument");
    IMBA_CHECK_ARG(type == IMBA_CACHE_BVH2 || type == IMBA_CACHE_BVH4, "unknown block type");
    IMBA_CHECK_ARG(node_count >= 0 && vec4_count >= 0, "negative count");
    bool exists;
    {
        std::ifstream in(path, std::ios::binary);
        exists = in && has_magic(in);
        uint64_t sz;
        if (exists && locate(in, (uint32_t)type, &sz))
            return fail(IMBA_ERR_INVALID_ARG, "%s already holds a block of type %d", path, type);
    }
    const size_t nb = node_bytes(type);
    std::vector<char> nbuf((const char*)nodes, (const char*)nodes + nb * node_count);
    std::vector<ImbaVec4> tbuf(tris, tris + vec4_count);
    if (type == IMBA_CACHE_BVH2) rebase_bvh2((ImbaNode*)nbuf.data(), node_count, -node_offset, -tris_offset);
    else rebase_bvh4((ImbaNode4*)nbuf.data(), node_count, -node_offset, -tris_offset);
    int r = type == IMBA_CACHE_BVH2 ? rebase_ids_bvh2(tbuf.data(), vec4_count, -tri_id_offset)
                                    : rebase_ids_bvh4(tbuf.data(), vec4_count, -tri_id_offset);
    if (r) return r;
    std::ofstream out(path, exists ? std::ios::binary | std::ios::app : std::ios::binary | std::ios::trunc);
    if (!out) return fail(IMBA_ERR_INVALID_ARG, "cannot open %s for writing", path);
    if (!exists) out.write((const char*)&kMagic, 4);
    const uint32_t t = (uint32_t)type, nc = (uint32_t)node_count, pc = (uint32_t)vec4_count;
    const uint64_t size = 4 + 8 + nb * (uint64_t)node_count + sizeof(ImbaVec4) * (uint64_t)vec4_count;
    out.write((const char*)&size, 8);
    out.write((const char*)&t, 4);
    out.write((const char*)&nc, 4);
    out.write((const char*)&pc, 4);
    out.write(nbuf.data(), (std::streamsize)nbuf.size());
    out.write((const char*)tbuf.data(), (std::streamsize)(sizeof(ImbaVec4) * tbuf.size()));
    if (!out) return fail(IMBA_ERR_INVALID_ARG, "write to %s failed", path);
    return IMBA_OK;
}

int cache_load(const char* path, int type, int node_offset, int tris_offset, int tri_id_offset,
               std::vector<char>& nodes, std::vector<ImbaVec4>& tris) {
    IMBA_CHECK_ARG(path, "NULL path");
    IMBA_CHECK_ARG(type == IMBA_CACHE_BVH2 || type == IMBA_CACHE_BVH4, "unknown block type");
    std::ifstream in(path, std::ios::binary);
    if (!in) return fail(IMBA_ERR_INVALID_ARG, "cannot open %s", path);
    if (!has_magic(in)) return fail(IMBA_ERR_BAD_BVH, "%s is not an acceleration cache file", path);
    uint64_t size = 0;
    if (!locate(in, (uint32_t)type, &size)) return fail(IMBA_ERR_BAD_BVH, "%s has no block of type %d", path, type);
    uint32_t nc = 0, pc = 0;
    if (!in.read((char*)&nc, 4) || !in.read((char*)&pc, 4)) return fail(IMBA_ERR_BAD_BVH, "%s: truncated block", path);
    const size_t nb = node_bytes(type);
    if (size != 4 + 8 + nb * (uint64_t)nc + sizeof(ImbaVec4) * (uint64_t)pc || nc > (1u << 30) || pc > (1u << 30))
        return fail(IMBA_ERR_BAD_BVH, "%s: block size %llu does not match %u nodes and %u records", path,
                    (unsigned long long)size, nc, pc);
    nodes.resize(nb * nc);
    tris.resize(pc);
    if (!in.read(nodes.data(), (std::streamsize)nodes.size()) ||
        !in.read((char*)tris.data(), (std::streamsize)(sizeof(ImbaVec4) * pc)))
        return fail(IMBA_ERR_BAD_BVH, "%s: truncated block", path);
    if (type == IMBA_CACHE_BVH2) rebase_bvh2((ImbaNode*)nodes.data(), (int)nc, node_offset, tris_offset);
    else rebase_bvh4((ImbaNode4*)nodes.data(), (int)nc, node_offset, tris_offset);
    return type == IMBA_CACHE_BVH2 ? rebase_ids_bvh2(tris.data(), (int)pc, tri_id_offset)
                                   : rebase_ids_bvh4(tris.data(), (int)pc, tri_id_offset);
}

}  // namespace imba

using namespace imba;

extern "C" {

int imba_bvh_store(const char* path, int type, const void* nodes, int node_count, const ImbaVec4* tris,
                   int vec4_count, int node_offset, int tris_offset, int tri_id_offset) {
    return cache_store(path, type, nodes, node_count, tris, vec4_count, node_offset, tris_offset, tri_id_offset);
}

int imba_bvh_load(const char* path, int type, int node_offset, int tris_offset, int tri_id_offset, void* nodes_out,
                  int node_cap, ImbaVec4* tris_out, int vec4_cap, int* node_count, int* vec4_count) {
    IMBA_CHECK_ARG(node_count && vec4_count, "NULL count argument");
    std::vector<char> nodes;
    std::vector<ImbaVec4> tris;
    int r = cache_load(path, type, node_offset, tris_offset, tri_id_offset, nodes, tris);
    if (r) return r;
    const int nc = (int)(nodes.size() / node_bytes(type)), pc = (int)tris.size();
    *node_count = nc;
    *vec4_count = pc;
    if (!nodes_out || !tris_out || node_cap < nc || vec4_cap < pc)
        return fail(IMBA_ERR_CAPACITY, "need room for %d nodes and %d Vec4 records", nc, pc);
    std::memcpy(nodes_out, nodes.data(), nodes.size());
    std::memcpy(tris_out, tris.data(), sizeof(ImbaVec4) * tris.size());
    return IMBA_OK;
}

int imba_bvh_has_block(const char* path, int type) { return cache_has_block(path, type); }

}  // extern "C"
