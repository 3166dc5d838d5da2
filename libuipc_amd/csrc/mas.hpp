// MAS hierarchy tables shared by the host build (mas.hip) and the device
// build (mas_init.hip).  BANKSIZE = 16 nodes per cluster, at most 6 levels
// (finite_element/mas_preconditioner_engine.h:27-45).
#pragma once
#include <vector>

#include "linsys.hpp"

namespace mi
{
constexpr int kMasBank   = 16;
constexpr int kMasMaxLvl = 6;

inline int mas_bank_align(int n)
{
    return (n + kMasBank - 1) / kMasBank * kMasBank;
}

// MASPreconditionerEngine::compute_num_levels (mas_preconditioner_engine.cu:1958-1972)
inline int mas_num_levels(int vert_num)
{
    int n_level = 1, sz = mas_bank_align(vert_num);
    while(sz > kMasBank)
    {
        sz /= kMasBank;
        n_level++;
        sz = mas_bank_align(sz);
    }
    n_level++;
    return n_level < kMasMaxLvl ? n_level : kMasMaxLvl;
}

// The tables of init_neighbor / init_matrix (mas_preconditioner_engine.cu:1974-2078):
// partition slots, going_next over all levels, level sizes and the per-vertex
// level tables.  fine_connect and prefix_orig are working state of the host
// build only.
struct MasHierarchy
{
    int32_t               V = 0, N = 0, L = 0, total = 0, level1_begin = 0, padded = 0;
    std::vector<int32_t>  part_to_real, real_to_part, going_next, prefix_orig;
    std::vector<uint32_t> fine_connect;
    std::vector<int32_t>  level_count, level_begin;  // [0..L]
    std::vector<int32_t>  table;                     // V * kMasMaxLvl
    bool                  has_unpart = false;
};

// Vertex adjacency on the device: sorted, unique neighbours of every vertex
// over the pairs of the given simplices (device arrays, any may be NULL;
// pairs with an end outside [0, V) or equal ends are skipped), as CSR.
struct DeviceAdjacency
{
    DBuf<uint32_t> nstart, nnum, nlist;  // V, V, M
    int64_t        M = 0;
};
void build_adjacency_device(int32_t          V,
                            const int32_t*   tets,
                            int64_t          n_tets,
                            const int32_t*   tris,
                            int64_t          n_tris,
                            const int32_t*   edges,
                            int64_t          n_edges,
                            DeviceAdjacency& out,
                            hipStream_t      s);

// geometry::mesh_partition on the device: parts of at most K vertices, ids
// 0..P-1 written to `part` (device, V entries).  Returns P.
int32_t mesh_partition_device(int32_t        V,
                              const int32_t* tets,
                              int64_t        n_tets,
                              const int32_t* tris,
                              int64_t        n_tris,
                              const int32_t* edges,
                              int64_t        n_edges,
                              int32_t        K,
                              int32_t*       part,
                              hipStream_t    s);

// The same tables built by device kernels from the topology and partition of
// the descriptor (device copies where the descriptor gives them).
MasHierarchy mas_build_hierarchy_device(const mi_precond_desc& d);
}  // namespace mi
