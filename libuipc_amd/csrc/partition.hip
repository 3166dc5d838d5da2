// Mesh partition into parts of at most K vertices on the device (gfx950),
// standing where geometry::mesh_partition (src/geometry/mesh_partition.cpp:77-230)
// runs METIS on the host for the MAS preconditioner.
//
// Same contract as mesh_partition: adjacency from the highest-dimensional
// simplices given (tets, else triangles, else edges); no adjacency -> part
// i / K; at most one part (V <= K) -> connected-component labels numbered by
// the lowest vertex of each component; otherwise parts of at most K vertices
// with ids 0..P-1, none empty.  The k-way partition itself is a different
// algorithm from METIS (so different ids), built for a GPU:
//   match    rounds of handshake matching on the cluster graph: every cluster
//            proposes to the neighbour cluster of highest score
//            w (1/size_c + 1/size_d) (w = fine edges between them) that keeps the
//            merged size <= K, equal scores ranked by a symmetric hash of the
//            pair (then the lower id); mutual proposals merge into the lower
//            id.  The cluster graph of a round is the fine
//            adjacency relabelled, sorted (radix) and run-length counted.
//            Scores are combined with 64-bit atomicMax, so the result does
//            not depend on scheduling; rounds end when nothing merges.
//   pack     clusters in order of their lowest vertex are packed greedily
//            into parts of at most K vertices (a sequential scan over the
//            cluster sizes on the host): banks of a part that are not
//            connected are block-diagonal in the MAS local solve, so packing
//            only removes padding.
#include <algorithm>
#include <vector>

#include "mas.hpp"

namespace mi
{
namespace
{
constexpr int kT = 256;

int grid_p(int64_t n)
{
    return std::max(1, std::min(ceil_div(n, kT), device_cu_count() * 16));
}

#define MI_GRID_STRIDE(i, n) \
    for(int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < (n); i += (int64_t)gridDim.x * blockDim.x)

__global__ void iota_kernel(int32_t* p, int32_t n)
{
    MI_GRID_STRIDE(i, n)
    p[i] = (int32_t)i;
}

__global__ void ones_kernel(int32_t* p, int32_t n)
{
    MI_GRID_STRIDE(i, n)
    p[i] = 1;
}

__global__ void reset_best_kernel(unsigned long long* best, uint32_t* target, int64_t n)
{
    MI_GRID_STRIDE(i, n)
    {
        best[i]   = 0ull;
        target[i] = 0xFFFFFFFFu;
    }
}

__global__ void seq_part_kernel(int32_t* part, int32_t V, int32_t K)
{
    MI_GRID_STRIDE(i, V)
    part[i] = (int32_t)(i / K);
}

// cluster pair of every directed fine edge, key c * V + d (c == d: sentinel V * V)
__global__ void cpair_kernel(int32_t V,
                             const uint32_t* __restrict__ nstart,
                             const uint32_t* __restrict__ nnum,
                             const uint32_t* __restrict__ nlist,
                             const int32_t* __restrict__ cv,
                             uint64_t* keys)
{
    const uint64_t VV = (uint64_t)V * (uint64_t)V;
    MI_GRID_STRIDE(a, V)
    {
        const uint64_t c = (uint64_t)cv[a];
        const uint32_t b = nstart[a], e = b + nnum[a];
        for(uint32_t q = b; q < e; ++q)
        {
            const uint64_t d = (uint64_t)cv[nlist[q]];
            keys[q]          = c != d ? c * (uint64_t)V + d : VV;
        }
    }
}

// symmetric pair hash: both ends of a cluster edge rank it the same, so that
// equal scores do not chain the proposals along vertex order
__device__ __forceinline__ uint32_t pair_hash(uint32_t c, uint32_t d, uint32_t seed)
{
    uint64_t x = ((uint64_t)min(c, d) << 32 | (uint64_t)max(c, d)) + 0x9E3779B97F4A7C15ull * (seed + 1);
    x ^= x >> 30;
    x *= 0xBF58476D1CE4E5B9ull;
    x ^= x >> 27;
    x *= 0x94D049BB133111EBull;
    x ^= x >> 31;
    return (uint32_t)x;
}

// the run of equal (c, d) keys at i (its head), or false
__device__ __forceinline__ bool cluster_edge(const uint64_t* keys,
                                             int64_t         i,
                                             int64_t         M,
                                             int32_t         V,
                                             const int32_t*  sz,
                                             int32_t         K,
                                             uint32_t        seed,
                                             uint32_t&       c,
                                             uint32_t&       d,
                                             unsigned long long& packed)
{
    const uint64_t VV = (uint64_t)V * (uint64_t)V;
    const uint64_t k  = keys[i];
    if(k >= VV || (i > 0 && keys[i - 1] == k))
        return false;
    int64_t w = 1;
    while(i + w < M && keys[i + w] == k)
        ++w;
    c                = (uint32_t)(k / (uint64_t)V);
    d                = (uint32_t)(k % (uint64_t)V);
    const int32_t sc = sz[c], sd = sz[d];
    if(sc + sd > K)
        return false;
    const float score = (float)w * ((float)(sc + sd) / ((float)sc * (float)sd));
    packed            = ((unsigned long long)__float_as_uint(score) << 32) | pair_hash(c, d, seed);
    return true;
}

// best[c] = max over c's cluster edges of (score, pair hash); the score
// favours heavy edges and small clusters, which keeps the sizes even
__global__ void cbest_kernel(const uint64_t* __restrict__ keys,
                             int64_t M,
                             int32_t V,
                             const int32_t* __restrict__ sz,
                             int32_t             K,
                             uint32_t            seed,
                             unsigned long long* best)
{
    MI_GRID_STRIDE(i, M)
    {
        uint32_t           c, d;
        unsigned long long p;
        if(cluster_edge(keys, i, M, V, sz, K, seed, c, d, p))
            atomicMax(&best[c], p);
    }
}

// target[c] = the lowest d whose edge reaches best[c]
__global__ void ctarget_kernel(const uint64_t* __restrict__ keys,
                               int64_t M,
                               int32_t V,
                               const int32_t* __restrict__ sz,
                               int32_t K,
                               uint32_t seed,
                               const unsigned long long* __restrict__ best,
                               uint32_t* target)
{
    MI_GRID_STRIDE(i, M)
    {
        uint32_t           c, d;
        unsigned long long p;
        if(cluster_edge(keys, i, M, V, sz, K, seed, c, d, p) && p == best[c])
            atomicMin(&target[c], d);
    }
}

// mutual proposals merge into the lower cluster id
__global__ void cmatch_kernel(int32_t V,
                              const int32_t* __restrict__ sz,
                              const uint32_t* __restrict__ target,
                              int32_t*  sz_new,
                              int32_t*  redirect,
                              uint32_t* merges)
{
    MI_GRID_STRIDE(c64, V)
    {
        const int32_t c = (int32_t)c64;
        const int32_t s = sz[c];
        redirect[c]     = c;
        sz_new[c]       = s;
        if(s == 0)
            continue;
        const uint32_t d = target[c];
        if(d >= (uint32_t)V || target[d] != (uint32_t)c)
            continue;
        if((uint32_t)c < d)
        {
            sz_new[c] = s + sz[d];
            atomicAdd(merges, 1u);
        }
        else
        {
            sz_new[c]   = 0;
            redirect[c] = (int32_t)d;
        }
    }
}

__global__ void relabel_kernel(int32_t V, const int32_t* __restrict__ redirect, int32_t* cv)
{
    MI_GRID_STRIDE(v, V)
    cv[v] = redirect[cv[v]];
}

__global__ void map_kernel(int32_t V, const int32_t* __restrict__ cv, const int32_t* __restrict__ part_of,
                           int32_t* part)
{
    MI_GRID_STRIDE(v, V)
    part[v] = part_of[cv[v]];
}

// connected components: label = lowest vertex of the component, by min-label
// propagation over the adjacency plus pointer jumping
__global__ void cc_step_kernel(int32_t V,
                               const uint32_t* __restrict__ nstart,
                               const uint32_t* __restrict__ nnum,
                               const uint32_t* __restrict__ nlist,
                               int32_t*  lab,
                               uint32_t* changed)
{
    MI_GRID_STRIDE(a, V)
    {
        int32_t        m = lab[a];
        const uint32_t b = nstart[a], e = b + nnum[a];
        for(uint32_t q = b; q < e; ++q)
            m = min(m, lab[nlist[q]]);
        m = min(m, lab[m]);
        if(m < lab[a])
        {
            atomicMin(&lab[a], m);
            atomicMin(&lab[lab[a]], m);
            changed[0] = 1u;
        }
    }
}

// roots (lab[v] == v) numbered in vertex order
__global__ void cc_root_kernel(int32_t V, const int32_t* __restrict__ lab, uint32_t* flag)
{
    MI_GRID_STRIDE(v, V)
    flag[v] = lab[v] == (int32_t)v ? 1u : 0u;
}

__global__ void cc_label_kernel(int32_t V, const int32_t* __restrict__ lab, const uint32_t* __restrict__ rank,
                                int32_t* part)
{
    MI_GRID_STRIDE(v, V)
    part[v] = (int32_t)rank[lab[v]];
}
}  // namespace

int32_t mesh_partition_device(int32_t        V,
                              const int32_t* tets,
                              int64_t        n_tets,
                              const int32_t* tris,
                              int64_t        n_tris,
                              const int32_t* edges,
                              int64_t        n_edges,
                              int32_t        K,
                              int32_t*       part,
                              hipStream_t    s)
{
    MI_REQUIRE(K > 0, MI_ERR_ARG, "mesh_partition: part_max_size must be greater than zero");
    MI_REQUIRE(V >= 0, MI_ERR_ARG, "mesh_partition: negative vertex count");
    if(V == 0)
        return 0;
    MI_REQUIRE(part != nullptr, MI_ERR_ARG, "mesh_partition: part_ids is NULL");
    // highest-dimensional simplices only (mesh_partition.cpp:35-57)
    DeviceAdjacency adj;
    if(tets && n_tets > 0)
        build_adjacency_device(V, tets, n_tets, nullptr, 0, nullptr, 0, adj, s);
    else if(tris && n_tris > 0)
        build_adjacency_device(V, nullptr, 0, tris, n_tris, nullptr, 0, adj, s);
    else
        build_adjacency_device(V, nullptr, 0, nullptr, 0, edges, edges ? n_edges : 0, adj, s);
    const int64_t M = adj.M;
    if(M == 0)
    {
        seq_part_kernel<<<grid_p(V), kT, 0, s>>>(part, V, K);
        MI_HIP(hipGetLastError());
        MI_HIP(hipStreamSynchronize(s));
        return (V - 1) / K + 1;
    }
    DBuf<uint32_t> flag, chunk;
    DBuf<int64_t>  tot;
    tot.resize(1);
    if(1 + (V - 1) / K <= 1)
    {
        // connected components (label_connected_vertices)
        DBuf<int32_t>  lab;
        DBuf<uint32_t> changed;
        lab.resize((size_t)V);
        changed.resize(1);
        iota_kernel<<<grid_p(V), kT, 0, s>>>(lab.data(), V);
        for(;;)
        {
            MI_HIP(hipMemsetAsync(changed.data(), 0, sizeof(uint32_t), s));
            cc_step_kernel<<<grid_p(V), kT, 0, s>>>(V, adj.nstart.data(), adj.nnum.data(), adj.nlist.data(),
                                                    lab.data(), changed.data());
            uint32_t c = 0;
            MI_HIP(hipMemcpyAsync(&c, changed.data(), sizeof(c), hipMemcpyDeviceToHost, s));
            MI_HIP(hipStreamSynchronize(s));
            if(!c)
                break;
        }
        flag.resize((size_t)V);
        cc_root_kernel<<<grid_p(V), kT, 0, s>>>(V, lab.data(), flag.data());
        exclusive_scan_u32(flag.data(), V, chunk, tot.data(), s);
        cc_label_kernel<<<grid_p(V), kT, 0, s>>>(V, lab.data(), flag.data(), part);
        int64_t n = 0;
        MI_HIP(hipMemcpyAsync(&n, tot.data(), sizeof(n), hipMemcpyDeviceToHost, s));
        MI_HIP(hipGetLastError());
        MI_HIP(hipStreamSynchronize(s));
        return (int32_t)n;
    }
    // matching rounds
    DBuf<int32_t>            cv, sz, sz_new, redirect;
    DBuf<unsigned long long> best;
    DBuf<uint64_t>           keys[2];
    DBuf<int32_t>            vals[2];
    DBuf<uint32_t>           hist, merges, target;
    DBuf<int64_t>            cnt;
    cv.resize((size_t)V);
    sz.resize((size_t)V);
    sz_new.resize((size_t)V);
    redirect.resize((size_t)V);
    best.resize((size_t)V);
    target.resize((size_t)V);
    keys[0].resize((size_t)M);
    keys[1].resize((size_t)M);
    vals[0].resize((size_t)M);
    vals[1].resize((size_t)M);
    merges.resize(1);
    cnt.resize(1);
    MI_HIP(hipMemcpyAsync(cnt.data(), &M, sizeof(M), hipMemcpyHostToDevice, s));
    iota_kernel<<<grid_p(V), kT, 0, s>>>(cv.data(), V);
    int32_t* szc = sz.data();      // cluster sizes, indexed by the lowest vertex
    int32_t* szn = sz_new.data();  // sizes after the round's merges
    ones_kernel<<<grid_p(V), kT, 0, s>>>(szc, V);
    const int bits = radix_key_bits((uint64_t)V * (uint64_t)V);
    for(int round = 0; round < 64; ++round)
    {
        cpair_kernel<<<grid_p(V), kT, 0, s>>>(V, adj.nstart.data(), adj.nnum.data(), adj.nlist.data(), cv.data(),
                                              keys[0].data());
        const int cur = radix_sort_pairs(keys[0].data(), vals[0].data(), keys[1].data(), vals[1].data(), cnt.data(),
                                         M, bits, hist, chunk, s);
        reset_best_kernel<<<grid_p(V), kT, 0, s>>>(best.data(), target.data(), V);
        cbest_kernel<<<grid_p(M), kT, 0, s>>>(keys[cur].data(), M, V, szc, K, (uint32_t)round, best.data());
        ctarget_kernel<<<grid_p(M), kT, 0, s>>>(keys[cur].data(), M, V, szc, K, (uint32_t)round, best.data(),
                                                target.data());
        MI_HIP(hipMemsetAsync(merges.data(), 0, sizeof(uint32_t), s));
        cmatch_kernel<<<grid_p(V), kT, 0, s>>>(V, szc, target.data(), szn, redirect.data(), merges.data());
        relabel_kernel<<<grid_p(V), kT, 0, s>>>(V, redirect.data(), cv.data());
        std::swap(szc, szn);
        uint32_t m = 0;
        MI_HIP(hipMemcpyAsync(&m, merges.data(), sizeof(m), hipMemcpyDeviceToHost, s));
        MI_HIP(hipStreamSynchronize(s));
        if(m == 0)
            break;
    }
    // pack clusters (in order of their lowest vertex) into parts of <= K
    std::vector<int32_t> hsz((size_t)V), part_of((size_t)V, -1);
    MI_HIP(hipMemcpyAsync(hsz.data(), szc, sizeof(int32_t) * V, hipMemcpyDeviceToHost, s));
    MI_HIP(hipStreamSynchronize(s));
    int32_t parts = 0, fill = K;
    for(int32_t c = 0; c < V; ++c)
    {
        if(hsz[c] == 0)
            continue;
        if(fill + hsz[c] > K)
        {
            ++parts;
            fill = 0;
        }
        fill += hsz[c];
        part_of[c] = parts - 1;
    }
    MI_HIP(hipMemcpyAsync(redirect.data(), part_of.data(), sizeof(int32_t) * V, hipMemcpyHostToDevice, s));
    map_kernel<<<grid_p(V), kT, 0, s>>>(V, cv.data(), redirect.data(), part);
    MI_HIP(hipGetLastError());
    MI_HIP(hipStreamSynchronize(s));
    return parts;
}
}  // namespace mi
