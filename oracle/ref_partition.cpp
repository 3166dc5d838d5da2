// TEST INFRASTRUCTURE ONLY.  The partition the reference would give a mesh:
// geometry::mesh_partition (src/geometry/mesh_partition.cpp:77-230) restated
// around the reference's own METIS 5.2.1 port (src/geometry/metis/, compiled
// in place by oracle/Makefile into oracle/_ref/libref_partition.so).  Tests
// use it to measure the device partitioner's quality against METIS; it is never
// linked into the product.
//
// Same control flow as mesh_partition:
//   * adjacency from the highest-dimensional simplices present (tets, else
//     triangles, else edges), sorted unique neighbours (build_adjacency, 17-74);
//   * no adjacency: part = i / part_max_size (106-112);
//   * one part or fewer: connected-component labels in order of the lowest
//     vertex of each component (label_connected_vertices, 117-130);
//   * METIS_PartGraphKway with default options and seed 0, retried with
//     block_size - 1 while a part exceeds part_max_size, sequential fallback
//     (132-192).
#include <metis.h>

#include <algorithm>
#include <cstdint>
#include <vector>

namespace
{
void adjacency(int32_t                 V,
               const int32_t*          simp,
               int64_t                 n,
               int                     k,
               std::vector<idx_t>&     xadj,
               std::vector<idx_t>&     adj)
{
    std::vector<std::vector<idx_t>> nb((size_t)V);
    for(int64_t t = 0; t < n; ++t)
        for(int i = 0; i < k; ++i)
            for(int j = i + 1; j < k; ++j)
            {
                int32_t a = simp[k * t + i], b = simp[k * t + j];
                if(a != b && a >= 0 && b >= 0 && a < V && b < V)
                {
                    nb[a].push_back(b);
                    nb[b].push_back(a);
                }
            }
    xadj.assign((size_t)V + 1, 0);
    adj.clear();
    for(int32_t v = 0; v < V; ++v)
    {
        auto& s = nb[v];
        std::sort(s.begin(), s.end());
        s.erase(std::unique(s.begin(), s.end()), s.end());
        adj.insert(adj.end(), s.begin(), s.end());
        xadj[v + 1] = (idx_t)adj.size();
    }
}
}  // namespace

extern "C" int ref_mesh_partition(int32_t        V,
                                  const int32_t* tets,
                                  int64_t        n_tets,
                                  const int32_t* tris,
                                  int64_t        n_tris,
                                  const int32_t* edges,
                                  int64_t        n_edges,
                                  int32_t        part_max,
                                  int32_t*       out)
{
    if(part_max <= 0)
        return -1;
    if(V <= 0)
        return 0;
    std::vector<idx_t> xadj, adj;
    if(tets && n_tets > 0)
        adjacency(V, tets, n_tets, 4, xadj, adj);
    else if(tris && n_tris > 0)
        adjacency(V, tris, n_tris, 3, xadj, adj);
    else if(edges && n_edges > 0)
        adjacency(V, edges, n_edges, 2, xadj, adj);
    else
        xadj.assign((size_t)V + 1, 0);
    if(adj.empty())
    {
        for(int32_t i = 0; i < V; ++i)
            out[i] = i / part_max;
        return 0;
    }
    int32_t block = part_max;
    idx_t   np    = (idx_t)(1 + (V - 1) / block);
    if(np <= 1)
    {
        std::vector<int32_t> label((size_t)V, -1), queue;
        int32_t              next = 0;
        for(int32_t s = 0; s < V; ++s)
        {
            if(label[s] >= 0)
                continue;
            label[s] = next;
            queue.assign(1, s);
            for(size_t q = 0; q < queue.size(); ++q)
                for(idx_t e = xadj[queue[q]]; e < xadj[queue[q] + 1]; ++e)
                    if(label[adj[e]] < 0)
                    {
                        label[adj[e]] = next;
                        queue.push_back(adj[e]);
                    }
            ++next;
        }
        std::copy(label.begin(), label.end(), out);
        return 0;
    }
    std::vector<idx_t> part((size_t)V, 0);
    bool               ok = false;
    while(np >= 2 && block >= 1)
    {
        idx_t cut = 0, ncon = 1, nv = V;
        idx_t opt[METIS_NOPTIONS];
        METIS_SetDefaultOptions(opt);
        opt[METIS_OPTION_SEED] = 0;
        if(METIS_PartGraphKway(&nv, &ncon, xadj.data(), adj.data(), nullptr, nullptr, nullptr, &np, nullptr,
                               nullptr, opt, &cut, part.data())
           != METIS_OK)
            break;
        std::vector<int32_t> size((size_t)np, 0);
        for(int32_t i = 0; i < V; ++i)
            ++size[part[i]];
        if(*std::max_element(size.begin(), size.end()) <= part_max)
        {
            ok = true;
            break;
        }
        if(block == 1)
            break;
        --block;
        np = (idx_t)(1 + (V - 1) / block);
    }
    for(int32_t i = 0; i < V; ++i)
        out[i] = ok ? (int32_t)part[i] : i / part_max;
    return 0;
}
