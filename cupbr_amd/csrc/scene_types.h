// Device-resident scene layout (HBM), shared by every kernel.
//
//   Node4    128-byte BVH4 node as built on the host: per-axis child boxes as
//            float4 SoA (96 B) + 4 child words (16 B) + 16 B spare.
//   NodeH    the same node as uploaded and traversed: binary16 plane offsets
//            (below), 128-byte aligned so one node is one HBM/L2 line.
//   PrimRec  48-byte primitive record in BVH leaf order:
//              triangle  v1.xyz | e1 = v2-v1 | e2 = v3-v1  (precomputed edges give the
//                        same FP32 values the reference computes per test,
//                        TriangleDetail.h:27-28)
//              sphere    center.xyz | radius
//            + reference rank, object id, kind/flags in the last 3 words.
//   UnbRec   planes and quads (infinite reference boxes), tested for every ray.
//   ObjInfo  per scene object: the box the reference tests before the object
//            (finite only for spheres and meshes), type, rank base.
//
// Child word: 0xFFFFFFFF empty; bit 31 set: leaf, prims [(w>>4)&0x7FFFFFF, +((w&15)+1));
// otherwise the index of an inner node.
#pragma once

#include "../../include/cupbr.h"
#include "vecmath.h"

namespace cupbr {

constexpr uint32_t kChildEmpty = 0xFFFFFFFFu;
constexpr uint32_t kLeafBit = 0x80000000u;
constexpr uint32_t kMaxLeafPrims = 16;
constexpr uint32_t kNoHit = 0xFFFFFFFFu;   // closest-hit result: miss
constexpr uint32_t kUnbBit = 0x80000000u;  // closest-hit result: index into the unbounded list

CUPBR_HD uint32_t leaf_word(uint32_t first, uint32_t count) { return kLeafBit | (first << 4) | (count - 1); }
CUPBR_HD uint32_t leaf_first(uint32_t w) { return (w >> 4) & 0x7FFFFFFu; }
CUPBR_HD uint32_t leaf_count(uint32_t w) { return (w & 15u) + 1u; }

struct alignas(128) Node4 {
    float lo_x[4], lo_y[4], lo_z[4];
    float hi_x[4], hi_y[4], hi_z[4];
    uint32_t child[4];
    uint32_t spare[4];
};
static_assert(sizeof(Node4) == 128, "Node4 must be one 128-byte line");

// The same node as the device reads it: each child plane as a binary16 offset from a reference
// point c of the node (the middle of its children's finite extent, as a float), in units of 2^k
// per axis, rounded outward with a margin; empty slots and unbounded sides are binary16
// infinities. Per axis two
// 16-byte blocks, one per sign of the ray's inverse direction, each holding that sign's near
// planes of the 4 children then their far planes, so a lane loads its axis in one 16-byte load:
// five loads per node (corner, x, y, z, child words) instead of seven.
struct alignas(128) NodeH {
    float cx, cy, cz;     // reference point c
    uint32_t exps;        // k_x, k_y, k_z as signed bytes 0..2 (k <= 0)
    uint16_t x[2][8];     // [0]: lo0..lo3 hi0..hi3 (inverse direction >= 0), [1]: hi0..hi3 lo0..lo3
    uint16_t y[2][8];
    uint16_t z[2][8];
    uint32_t child[4];
};
static_assert(sizeof(NodeH) == 128, "NodeH must be one 128-byte line");

enum PrimKind : uint32_t { kPrimTriangle = 0, kPrimSphere = 1 };
constexpr uint32_t kPrimNeedsBoxCheck = 4u;  // object box is finite (sphere, mesh): test it lazily
// normals[prim].w (bits): object id in bits 0..30, bit 31 set for a sphere (whose normal is
// computed from the hit point, not stored)
constexpr uint32_t kNormalSphereBit = 0x80000000u;

struct alignas(16) PrimRec {
    float4 a;  // tri: v1.x v1.y v1.z e1.x | sphere: c.x c.y c.z r
    float4 b;  // tri: e1.y e1.z e2.x e2.y
    float4 c;  // tri: e2.z | rank | object | kind+flags (as bits)
};
static_assert(sizeof(PrimRec) == 48, "PrimRec is 48 bytes");

enum UnbKind : int32_t { kUnbPlane = 0, kUnbQuad = 1 };
struct alignas(16) UnbRec {
    float4 p;   // position.xyz, kind (bits)
    float4 n;   // normal.xyz, rank (bits)
    float4 e1;  // quad extend1.xyz, object (bits)
    float4 e2;  // quad extend2.xyz
};

struct ObjInfo {
    float bmin[3];
    float bmax[3];
    int32_t type;
    uint32_t rank_base;
    uint32_t finite_box;
    uint32_t rank_count;  // primitives of the object (triangles of a mesh, else 1)
    uint32_t first_triangle;  // mesh: index of its first triangle in the scene's triangle array
    uint32_t pad;
};

// Everything a kernel needs, passed by value.
struct DevScene {
    const NodeH* hnodes;   // the BVH4 (the FP32 Node4 form stays on the host)
    const PrimRec* prims;
    const float4* normals;  // per prim in leaf order: triangle normal; w = object id | kNormalSphereBit
    const UnbRec* unb;
    const ObjInfo* objs;
    const cupbr_material* mats;
    const cupbr_light* lights;
    const float* env;  // env_w * env_h * 3
    const uint32_t* rank_prim;  // reference rank -> prim index, kUnbBit|unbounded index, or 0xFFFFFFFF
    uint32_t n_nodes, n_prims, n_unb, n_objs, n_lights;
    uint32_t env_w, env_h;
    int32_t use_env;
    cupbr_volume volume;
};

}  // namespace cupbr
