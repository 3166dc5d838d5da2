// MAS (multi-level additive Schwarz) preconditioner on gfx950.
//
// Replaces FEMMASPreconditioner (finite_element/fem_mas_preconditioner.cu:157-469)
// and MASPreconditionerEngine (finite_element/mas_preconditioner_engine.cu).
// BANKSIZE = 16 nodes (48 DoFs) per cluster, at most 6 levels, upper-triangle
// cluster storage of 136 3x3 blocks (mas_preconditioner_engine.h:27-45).
//
// Init (once per topology): adjacency from the simplices, partition slots
// (part_to_real / real_to_part) and the level hierarchy (connect masks,
// closure inside each 16-bank, elected-lane prefix -> going_next, level sizes,
// per-vertex level tables), restating the reference's integer kernels
// (mas_preconditioner_engine.cu:57-466, 1958-2078) -- bit-identical tables.
// Built here in host C++, or by the device kernels of mas_init.hip
// (MI_MAS_BUILD_DEVICE), which give the same tables.
//
// Assembly (per Newton iteration, GPU):
//   cluster matrices at level L are the in-bank blocks of the Galerkin operator
//   R_L A R_L^T (the reference's scatter k1 + k2, 530-938).  Here:
//     walk   one thread per upper block: the level-0 in-bank store (each cluster
//            entry has exactly one source, so a plain store), then up the levels
//            adding the block where its two ancestors are different nodes of one
//            bank (FP64 atomics, wave-aggregated when few destinations repeat);
//            the walk stops where the ancestors merge;
//     push   level by level, the full sum of every group of nodes that share a
//            parent becomes the parent's diagonal block (plain store; a parent's
//            children all live in one bank).  This replaces the reference's
//            per-block diagonal atomics at every level above the merge, which
//            would all land on the few top-level blocks.
//   invert one wave per cluster, FP64 symmetric Gauss-Jordan sweep
//          (the sweep of 1077-1175, in double); zero diagonal -> 1; stored as
//          FP64 or FP32 upper blocks.
//
// Apply (per PCG iteration, GPU; 2887-2960):
//   restrict level by level, the lowest node of every in-bank sibling group
//            stores the group's sum into the parent (level 0 gathers r through
//            part_to_real); the small top levels run in one workgroup;
//   coarse   one wave per coarse cluster, local solve Z = C^-1 R;
//   fine     one wave per fine cluster: gather r, local solve, and the
//            prolongation z_i = Z0_i + sum_L Z_L[ancestor_L(i)] (injection),
//            written straight to z.
//   The local solve reads the inverse as 9 coalesced planes of 136 blocks and
//   sums each row in a fixed order.
//   No global atomics in the apply.
#include <algorithm>
#include <cstring>
#include <numeric>
#include <type_traits>

#include <cstdlib>

#include "mas.hpp"

namespace mi
{
namespace
{
constexpr int kBank   = kMasBank;
constexpr int kSym    = kBank * (kBank + 1) / 2;  // 136
constexpr int kCm     = kSym * 9;                 // 1224 scalars per cluster
constexpr int kMaxLvl = kMasMaxLvl;
// the MAS restriction kernels run up to the first level with at most this many
// nodes; coarse clusters above it sum their nodes' descendants at that level
constexpr int32_t kTailLevel = 4096;

__host__ __device__ __forceinline__ int sym_index(int row, int col)
{
    return kBank * row - row * (row + 1) / 2 + col;
}
// Cluster inverses are stored as 9 planes of the 136 upper blocks:
// inv[k * kCm + q * kSym + blk] = entry q (column-major, r + 3 c) of block blk,
// so a wave reading one plane for 64 consecutive blocks is one coalesced load.
__host__ __device__ __forceinline__ int inv_at(int blk, int q)
{
    return q * kSym + blk;
}

inline int bank_align(int n)
{
    return mas_bank_align(n);
}
inline int popc(uint32_t x)
{
    return __builtin_popcount(x);
}
inline int ffs1(uint32_t x)
{
    return __builtin_ffs((int)x);
}
inline uint32_t lt_mask(int lane)
{
    return (1u << lane) - 1u;
}

// ------------------------------------------------------------------ host init
using Hierarchy = MasHierarchy;

uint32_t bank_closure(const uint32_t cache[kBank], int lane, uint32_t start)
{
    uint32_t conn = start, visited = 1u << lane;
    for(;;)
    {
        uint32_t todo = visited ^ conn;
        if(!todo)
            break;
        int nv = ffs1(todo) - 1;
        visited |= 1u << nv;
        conn |= cache[nv];
    }
    return conn;
}

Hierarchy build_hierarchy(const mi_precond_desc& d)
{
    Hierarchy h;
    const int32_t V = d.vert_num;
    h.V             = V;
    // adjacency, sorted unique per vertex (fem_mas_preconditioner.cu:163-203)
    std::vector<std::vector<uint32_t>> nb((size_t)V);
    auto add = [&](int32_t a, int32_t b)
    {
        if(a != b && a >= 0 && b >= 0 && a < V && b < V)
        {
            nb[a].push_back((uint32_t)b);
            nb[b].push_back((uint32_t)a);
        }
    };
    for(int64_t t = 0; t < d.n_tets; ++t)
        for(int i = 0; i < 4; ++i)
            for(int j = i + 1; j < 4; ++j)
                add(d.tets_host[4 * t + i], d.tets_host[4 * t + j]);
    for(int64_t t = 0; t < d.n_tris; ++t)
        for(int i = 0; i < 3; ++i)
            for(int j = i + 1; j < 3; ++j)
                add(d.tris_host[3 * t + i], d.tris_host[3 * t + j]);
    for(int64_t t = 0; t < d.n_edges; ++t)
        add(d.edges_host[2 * t], d.edges_host[2 * t + 1]);
    std::vector<uint32_t> nstart((size_t)V), nnum((size_t)V), nlist;
    for(int32_t v = 0; v < V; ++v)
    {
        auto& s = nb[v];
        std::sort(s.begin(), s.end());
        s.erase(std::unique(s.begin(), s.end()), s.end());
        nstart[v] = (uint32_t)nlist.size();
        nnum[v]   = (uint32_t)s.size();
        nlist.insert(nlist.end(), s.begin(), s.end());
        std::vector<uint32_t>().swap(s);
    }
    // partition slots (fem_mas_preconditioner.cu:291-356)
    int32_t max_part = 0;
    for(int32_t v = 0; v < V; ++v)
        max_part = std::max(max_part, d.part_ids_host[v]);
    std::vector<int32_t> psize((size_t)max_part + 1, 0), poff((size_t)max_part + 1, 0), pfill((size_t)max_part + 1, 0);
    for(int32_t v = 0; v < V; ++v)
        if(d.part_ids_host[v] >= 0)
            MI_REQUIRE(++psize[d.part_ids_host[v]] <= kBank, MI_ERR_ARG,
                       "MAS: a partition holds more than 16 vertices");
    int32_t N = 0;
    for(int32_t p = 0; p <= max_part; ++p)
    {
        poff[p] = N;
        N += bank_align(psize[p]);
    }
    h.N = N;
    h.part_to_real.assign((size_t)std::max(N, 1), -1);
    h.real_to_part.assign((size_t)V, -1);
    for(int32_t v = 0; v < V; ++v)
    {
        int32_t p = d.part_ids_host[v];
        if(p < 0)
        {
            h.has_unpart = true;
            continue;
        }
        int32_t s          = poff[p] + pfill[p]++;
        h.part_to_real[s]  = v;
        h.real_to_part[v]  = s;
    }
    // init_neighbor (1974-2024)
    const int max_nodes = std::max(N, V);
    h.padded            = bank_align(max_nodes);
    h.L                 = mas_num_levels(max_nodes);
    const int L         = h.L;
    h.going_next.assign((size_t)h.padded * L, -1);
    std::vector<int32_t> coarse((size_t)V * L, -1);
    h.table.assign((size_t)V * kMaxLvl, -1);
    h.fine_connect.assign((size_t)max_nodes, 0u);
    const int nbanks = N / kBank;
    h.prefix_orig.assign((size_t)std::max(nbanks, 1), 0);
    h.level_count.assign((size_t)L + 1, 0);
    h.level_begin.assign((size_t)L + 1, 0);
    if(N < 1)
        return h;
    // build_connect_mask_L0 (57-106)
    for(int32_t tid = 0; tid < N; ++tid)
    {
        int32_t idx = h.part_to_real[tid];
        if(idx < 0)
            continue;
        int      bank = tid / kBank;
        uint32_t msk  = 1u << (tid % kBank), nk = 0, st = nstart[idx];
        for(uint32_t q = 0; q < nnum[idx]; ++q)
        {
            uint32_t w  = nlist[st + q];
            int32_t  np = h.real_to_part[w];
            if(np < 0)
                continue;
            if(np / kBank == bank)
                msk |= 1u << (np % kBank);
            else
                nlist[st + nk++] = w;
        }
        nnum[idx]            = nk;
        h.fine_connect[idx] = msk;
    }
    // prepare_prefix_sum_L0 (108-156)
    for(int b = 0; b < nbanks; ++b)
    {
        uint32_t cache[kBank];
        for(int l = 0; l < kBank; ++l)
        {
            int32_t idx = h.part_to_real[b * kBank + l];
            cache[l]    = idx >= 0 ? h.fine_connect[idx] : 0u;
        }
        int cnt = 0;
        for(int l = 0; l < kBank; ++l)
        {
            int32_t idx = h.part_to_real[b * kBank + l];
            if(idx < 0)
                continue;
            uint32_t c           = bank_closure(cache, l, cache[l]);
            h.fine_connect[idx] = c;
            cnt += popc(c & lt_mask(l)) == 0;
        }
        h.prefix_orig[b] = cnt;
    }
    // build_level1 (158-212; host 1926-1972)
    std::vector<int32_t> psum((size_t)nbanks + 1, 0);
    for(int b = 0; b < nbanks; ++b)
        psum[b + 1] = psum[b] + h.prefix_orig[b];
    const int level1_begin = bank_align(std::max(bank_align(N), V));
    h.level1_begin         = level1_begin;
    for(int b = 0; b < nbanks; ++b)
    {
        uint32_t elected = 0;
        for(int l = 0; l < kBank; ++l)
        {
            int32_t idx = h.part_to_real[b * kBank + l];
            if(idx >= 0 && popc(h.fine_connect[idx] & lt_mask(l)) == 0)
                elected |= 1u << l;
        }
        for(int l = 0; l < kBank; ++l)
        {
            int32_t idx = h.part_to_real[b * kBank + l];
            if(idx < 0)
                continue;
            int el            = ffs1(h.fine_connect[idx]) - 1;
            int prefix        = popc(elected & lt_mask(el)) + psum[b];
            coarse[idx]       = prefix;
            h.going_next[idx] = prefix + level1_begin;
        }
    }
    h.level_count[1] = psum[nbanks];
    h.level_begin[1] = level1_begin;
    std::vector<uint32_t> next_mask((size_t)max_nodes);
    std::vector<int32_t>  next_pref((size_t)max_nodes / kBank + 2), next_psum((size_t)max_nodes / kBank + 2);
    for(int level = 1; level < L; ++level)
    {
        std::fill(next_mask.begin(), next_mask.end(), 0u);
        // build_connect_mask_Lx (214-294)
        for(int32_t tid = 0; tid < N; ++tid)
        {
            int32_t idx = h.part_to_real[tid];
            if(idx < 0)
                continue;
            int32_t cidx = coarse[(size_t)(level - 1) * V + idx];
            if(cidx < 0 || cidx >= N)
                continue;
            uint32_t conn = 0, nk = 0, st = nstart[idx];
            for(uint32_t q = 0; q < nnum[idx]; ++q)
            {
                uint32_t w  = nlist[st + q];
                int32_t  cc = coarse[(size_t)(level - 1) * V + w];
                if(cc < 0 || cc >= N)
                {
                    nlist[st + nk++] = w;
                    continue;
                }
                if(cidx / kBank == cc / kBank)
                    conn |= 1u << (cc % kBank);
                else
                    nlist[st + nk++] = w;
            }
            nnum[idx] = nk;
            next_mask[cidx] |= conn;
        }
        const int32_t Nx = h.level_count[level], begin = h.level_begin[level];
        const int     nbx = (Nx + kBank - 1) / kBank;
        // next_level_cluster (296-338)
        for(int b = 0; b < nbx; ++b)
        {
            uint32_t cache[kBank];
            for(int l = 0; l < kBank; ++l)
            {
                int i    = b * kBank + l;
                cache[l] = i < Nx ? ((1u << l) | next_mask[i]) : 0u;
            }
            int cnt = 0;
            for(int l = 0; l < kBank; ++l)
            {
                int i = b * kBank + l;
                if(i >= Nx)
                    continue;
                uint32_t c   = bank_closure(cache, l, cache[l]);
                next_mask[i] = c;
                cnt += popc(c & lt_mask(l)) == 0;
            }
            next_pref[b] = cnt;
        }
        // prefix_sum_Lx (340-388)
        next_psum[0] = 0;
        for(int b = 0; b < nbx; ++b)
            next_psum[b + 1] = next_psum[b] + next_pref[b];
        const int next_begin = begin + bank_align(Nx);
        for(int b = 0; b < nbx; ++b)
        {
            uint32_t elected = 0;
            for(int l = 0; l < kBank; ++l)
            {
                int i = b * kBank + l;
                if(i < Nx && popc(next_mask[i] & lt_mask(l)) == 0)
                    elected |= 1u << l;
            }
            int32_t nv[kBank];
            for(int l = 0; l < kBank; ++l)
            {
                int i = b * kBank + l;
                if(i >= Nx)
                    continue;
                int el                   = ffs1(next_mask[i]) - 1;
                nv[l]                    = popc(elected & lt_mask(el)) + next_psum[b];
                h.going_next[begin + i] = nv[l] + next_begin;
            }
            for(int l = 0; l < kBank; ++l)
                if(b * kBank + l < Nx)
                    next_mask[b * kBank + l] = (uint32_t)nv[l];
        }
        h.level_count[level + 1] = next_psum[nbx];
        h.level_begin[level + 1] = next_begin;
        // compute_next_level (390-409)
        for(int32_t idx = 0; idx < V; ++idx)
        {
            int32_t nx = coarse[(size_t)(level - 1) * V + idx];
            coarse[(size_t)level * V + idx] = (nx < 0 || nx >= V) ? -1 : (int32_t)next_mask[nx];
        }
    }
    h.total = h.level_begin[L];
    // aggregation_kernel (411-466)
    for(int32_t idx = 0; idx < V; ++idx)
    {
        int32_t first = h.going_next[idx];
        if(first < 0)
            continue;
        int32_t* tab = &h.table[(size_t)kMaxLvl * idx];
        tab[0]       = first;
        int32_t cur  = first;
        for(int l = 1; l < L - 1; ++l)
        {
            cur    = h.going_next[cur];
            tab[l] = cur;
        }
    }
    return h;
}

// ------------------------------------------------------------------ assembly kernels
__global__ void fill_f64_kernel(double* p, int64_t n)
{
    for(int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0.0;
}

template <typename T>
__global__ void fill_kernel(T* p, int64_t n)
{
    for(int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = T(0);
}

// add v (9 values, column-major) to cluster block (a, b) of the upper storage,
// aggregating lanes of the wave that target the same block
__device__ __forceinline__ void agg_atomic_add(double* hess, int64_t dst, const double v[9])
{
    // up to 4 rounds of leader-based aggregation, the rest adds directly
    bool pending = dst >= 0;
#pragma unroll 1
    for(int round = 0; round < 4; ++round)
    {
        uint64_t act = __ballot(pending);
        if(act == 0)
            return;
        int      leader = __ffsll((long long)act) - 1;
        int64_t  ld     = __shfl(dst, leader, kWave);
        bool     mine   = pending && dst == ld;
        uint64_t peers  = __ballot(mine);
        if(__popcll(peers) == 1 && round > 0)
            break;  // destinations no longer repeat: plain atomics below
#pragma unroll
        for(int q = 0; q < 9; ++q)
        {
            // every lane is active here (the loop's exits are wave-uniform)
            const double s = wave_total(mine ? v[q] : 0.0);
            if(__lane_id() == leader)
                atomic_add_f64(hess + ld + q, s);
        }
        if(mine)
            pending = false;
    }
    if(pending)
#pragma unroll
        for(int q = 0; q < 9; ++q)
            atomic_add_f64(hess + dst + q, v[q]);
}

// One upper block (r, c) of the MAS's range (local rows, sr / sc their
// partition slots; live = a block to walk, every lane of the wave calls this):
// level 0 by a plain store (each cluster entry has this one source, scatter
// k1, mas_preconditioner_engine.cu:530-682), the coarse levels' off-diagonal
// in-bank contributions until the ancestors merge by wave-aggregated atomics
// -- level 1 into the caller's LDS copy of its level-1 cluster instead when
// l1 is given (lds_base: that cluster's first entry in hess).
__device__ __forceinline__ void walk_block(bool           live,
                                           int64_t        u,
                                           int32_t        r,
                                           int32_t        c,
                                           int32_t        sr,
                                           int32_t        sc,
                                           const double* __restrict__ vals,
                                           int32_t        V,
                                           const int32_t* __restrict__ anc,
                                           int            L,
                                           double*        hess,
                                           double*        l1,
                                           int64_t        lds_base)
{
    double H[9];
    if(live)
    {
#pragma unroll
        for(int q = 0; q < 9; ++q)
            H[q] = vals[upper_at(u, q)];
        if(sr / kBank == sc / kBank)
        {
            // level 0: the only source of this cluster entry (scatter k1, 530-682)
            double* dst;
            if(sc >= sr)
            {
                dst = hess + (int64_t)(sc / kBank) * kCm + 9 * sym_index(sr % kBank, sc % kBank);
#pragma unroll
                for(int q = 0; q < 9; ++q)
                    dst[q] = H[q];
            }
            else
            {
                dst = hess + (int64_t)(sc / kBank) * kCm + 9 * sym_index(sc % kBank, sr % kBank);
#pragma unroll
                for(int i = 0; i < 3; ++i)
#pragma unroll
                    for(int j = 0; j < 3; ++j)
                        dst[i + 3 * j] = H[j + 3 * i];
            }
        }
        if(r == c)
            live = false;  // merged at level 0
    }
    else
    {
#pragma unroll
        for(int q = 0; q < 9; ++q)
            H[q] = 0.0;
    }
    // coarse levels: off-diagonal in-bank contributions until the ancestors merge
    for(int l = 1; l < L; ++l)
    {
        int64_t dst = -1;
        double  v[9];
        if(live)
        {
            int32_t a = anc[(int64_t)(l - 1) * V + r];
            int32_t b = anc[(int64_t)(l - 1) * V + c];
            if(a < 0 || b < 0 || a == b)
                live = false;
            else if(a / kBank == b / kBank)
            {
                if(b > a)
                {
                    dst = (int64_t)(a / kBank) * kCm + 9 * sym_index(a % kBank, b % kBank);
#pragma unroll
                    for(int q = 0; q < 9; ++q)
                        v[q] = H[q];
                }
                else
                {
                    dst = (int64_t)(a / kBank) * kCm + 9 * sym_index(b % kBank, a % kBank);
#pragma unroll
                    for(int i = 0; i < 3; ++i)
#pragma unroll
                        for(int j = 0; j < 3; ++j)
                            v[i + 3 * j] = H[j + 3 * i];
                }
            }
        }
        if(dst < 0)
        {
#pragma unroll
            for(int q = 0; q < 9; ++q)
                v[q] = 0.0;
        }
        if(__ballot(live) == 0)
            break;
        if(l == 1 && l1)
        {
            // the workgroup's own level-1 cluster: LDS atomics (an entry
            // outside it, which the build's check rules out, goes to memory)
            const int64_t e = dst - lds_base;
            if(dst >= 0 && e >= 0 && e < kCm)
#pragma unroll
                for(int q = 0; q < 9; ++q)
                    atomicAdd(&l1[e + q], v[q]);
            else if(dst >= 0)
#pragma unroll
                for(int q = 0; q < 9; ++q)
                    atomic_add_f64(hess + dst + q, v[q]);
        }
        else
            agg_atomic_add(hess, dst, v);
    }
}

// one thread per upper block (see the file comment)
__global__ __launch_bounds__(256) void mas_walk_kernel(int64_t        U,
                                                      const int32_t* __restrict__ rows,
                                                      const int32_t* __restrict__ cols,
                                                      const double* __restrict__ vals,
                                                      int32_t        blk_off,
                                                      int32_t        V,
                                                      const int32_t* __restrict__ real_to_part,
                                                      const int32_t* __restrict__ anc,  // (L-1) x V
                                                      int            L,
                                                      double*        hess)
{
    const int64_t u    = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    bool          live = u < U;
    int32_t       r = 0, c = 0, sr = -1, sc = -1;
    if(live)
    {
        r = rows[u] - blk_off;
        c = cols[u] - blk_off;
        live = r >= 0 && r < V && c >= 0 && c < V;
    }
    if(live)
    {
        sr   = real_to_part[r];
        sc   = real_to_part[c];
        live = sr >= 0 && sc >= 0;
    }
    walk_block(live, u, r, c, sr, sc, vals, V, anc, L, hess, nullptr, 0);
}


// One round of the push for four clusters, one per wave (k; live: a
// cluster to do): the full sum of each group of its nodes that share a
// parent -> the parent's diagonal block.  A parent in the cluster whose
// entries start at hess[l1_base] is written to l1 (LDS) when l1 is given.
// Every thread of the workgroup calls this (it holds barriers).
__device__ __forceinline__ void push_round(int64_t        k,
                                           bool           live,
                                           const int32_t* __restrict__ parent,
                                           double*        hess,
                                           double (*S)[kBank][9],
                                           int32_t (*P)[kBank],
                                           double*        l1,
                                           int64_t        l1_base)
{
    const int w    = threadIdx.x / kWave;
    const int lane = threadIdx.x % kWave;
    if(lane < kBank)
    {
        P[w][lane] = live ? parent[k * kBank + lane] : -1;
#pragma unroll
        for(int q = 0; q < 9; ++q)
            S[w][lane][q] = 0.0;
    }
    __syncthreads();
    if(live)
    {
        for(int blk = lane; blk < kSym; blk += kWave)
        {
            // blk -> (a, b), a <= b
            int a = 0, rem = blk;
            while(rem >= kBank - a)
            {
                rem -= kBank - a;
                ++a;
            }
            int     b  = a + rem;
            int32_t pa = P[w][a];
            if(pa < 0 || pa != P[w][b])
                continue;
            int rep = 0;
            while(P[w][rep] != pa)
                ++rep;
            const double* M = hess + k * kCm + 9 * blk;
            for(int i = 0; i < 3; ++i)
                for(int j = 0; j < 3; ++j)
                {
                    double val = M[i + 3 * j];
                    if(a != b)
                        val += M[j + 3 * i];
                    atomicAdd(&S[w][rep][i + 3 * j], val);
                }
        }
    }
    __syncthreads();
    if(live && lane < kBank)
    {
        int32_t pa = P[w][lane];
        bool    first = pa >= 0;
        for(int q = 0; q < lane && first; ++q)
            first = P[w][q] != pa;
        if(first)
        {
            const int64_t d = (int64_t)(pa / kBank) * kCm + 9 * sym_index(pa % kBank, pa % kBank);
            double*       D = (l1 && d >= l1_base && d < l1_base + kCm) ? l1 + (d - l1_base) : hess + d;
#pragma unroll
            for(int q = 0; q < 9; ++q)
                D[q] = S[w][lane][q];
        }
    }
    __syncthreads();  // S and P are reused by the next round
}

// push(level): for every level-`level` cluster (one wave each), the full sum of
// each group of its nodes that share a parent -> the parent's diagonal block
__global__ __launch_bounds__(256) void mas_push_kernel(int64_t        c_first,
                                                      int64_t        c_count,
                                                      const int32_t* __restrict__ parent,
                                                      double*        hess)
{
    __shared__ double  S[4][kBank][9];
    __shared__ int32_t P[4][kBank];
    const int     w    = threadIdx.x / kWave;
    const int64_t k    = c_first + (int64_t)blockIdx.x * 4 + w;
    const bool    live = (int64_t)blockIdx.x * 4 + w < c_count;
    push_round(k, live, parent, hess, S, P, nullptr, 0);
}

// The same walk in the partition order, one workgroup per 256-slot range
// whose parts make up one level-1 bank (the build checks this): every
// contribution to that level-1 cluster comes from the range's own rows, so
// the workgroup adds them in LDS and stores the cluster once, where the
// block-order walk spent most of its time in global atomics at level 1.
// Blocks are taken row by row (the upper blocks of each slot's vertex), each
// exactly once, as in the block-order walk; the level-0 push of the range's
// parts follows, into the same LDS cluster (the push kernel then starts at
// level 1).
__global__ __launch_bounds__(256) void mas_walk_range_kernel(const int32_t* __restrict__ row_ptr,  // upper BCOO, global rows
                                                            const int32_t* __restrict__ cols,
                                                            const double* __restrict__ vals,
                                                            int32_t        blk_off,
                                                            int32_t        V,
                                                            const int32_t* __restrict__ part_to_real,
                                                            const int32_t* __restrict__ real_to_part,
                                                            const int32_t* __restrict__ anc,
                                                            int            L,
                                                            const int32_t* __restrict__ range_bank,
                                                            int64_t        c1_first,
                                                            int32_t        nslots,
                                                            const int32_t* __restrict__ parent,
                                                            double*        hess)
{
    __shared__ double   l1[kCm];
    __shared__ uint32_t off[256 + 1];
    __shared__ uint32_t lds[256 / kWave + 1];
    const int     t     = threadIdx.x;
    const int32_t bank  = range_bank[blockIdx.x];
    const int64_t base1 = (c1_first + (bank < 0 ? 0 : bank)) * kCm;
    for(int k = t; k < kCm; k += 256)
        l1[k] = 0.0;
    // the range's rows and their upper-block counts
    const int32_t slot = blockIdx.x * 256 + t;
    const int32_t r    = slot < nslots ? part_to_real[slot] : -1;
    const int32_t b0   = r >= 0 ? row_ptr[blk_off + r] : 0;
    const uint32_t cnt = r >= 0 ? (uint32_t)(row_ptr[blk_off + r + 1] - b0) : 0u;
    uint32_t      total;
    const uint32_t o = block_exclusive_scan<256>(cnt, lds, &total);  // its barriers also order the l1 zeros
    off[t]           = o;
    if(t == 0)
        off[256] = total;
    __shared__ int32_t rs[256], bs[256];
    rs[t] = r;
    bs[t] = b0;
    __syncthreads();
    for(uint32_t i0 = 0; i0 < total; i0 += 256)
    {
        const uint32_t i    = i0 + t;
        bool           live = i < total;
        int32_t        rr = 0, c = 0, sr = -1, sc = -1;
        int64_t        u = 0;
        if(live)
        {
            // the slot holding flattened block i: the last offset <= i
            int lo = 0, hi = 255;
            while(lo < hi)
            {
                const int mid = (lo + hi + 1) >> 1;
                if(off[mid] <= i)
                    lo = mid;
                else
                    hi = mid - 1;
            }
            rr = rs[lo];
            u  = (int64_t)bs[lo] + (i - off[lo]);
            c  = cols[u] - blk_off;
            live = c >= 0 && c < V;
        }
        if(live)
        {
            sr   = real_to_part[rr];
            sc   = real_to_part[c];
            live = sr >= 0 && sc >= 0;
        }
        walk_block(live, u, rr, c, sr, sc, vals, V, anc, L, hess, l1, base1);
    }
    // the level-0 push of the range's 16 parts (whose clusters this workgroup
    // has just stored; the barriers order those stores before the reads):
    // each level-1 node's diagonal block, into the LDS cluster
    __shared__ double  S[4][kBank][9];
    __shared__ int32_t P[4][kBank];
    __syncthreads();
    const int64_t c0 = nslots / kBank;  // level-0 clusters
    for(int round = 0; round < 4; ++round)
    {
        const int64_t k = (int64_t)blockIdx.x * kBank + round * 4 + t / kWave;
        push_round(k, k < c0, parent, hess, S, P, l1, base1);
    }
    if(bank >= 0)
        for(int k = t; k < kCm; k += 256)
            hess[base1 + k] = l1[k];
}

// upper block index -> (row node a, column node b), a <= b (inverse of sym_index)
__device__ __forceinline__ void sym_rc(int blk, int& a, int& b)
{
    int r = 0;
#pragma unroll
    for(int t = 1; t < kBank; ++t)
        r += blk >= kBank * t - t * (t - 1) / 2 ? 1 : 0;
    a = r;
    b = blk - (kBank * r - r * (r - 1) / 2) + r;
}

// one wave per cluster: symmetric Gauss-Jordan sweep in FP64 (lane = column;
// the sweep of mas_preconditioner_engine.cu:1077-1175, in double).  The column
// is kept rotated so that the current pivot row is always register 0: after
// each pivot the registers shift down by one and the pivot's new value goes to
// the end, so every register index is static without unrolling the 48 pivots
// (a fully unrolled sweep spills).  After 48 pivots the rotation is the
// identity and the registers hold -A^{-1}.
template <typename TO>
__global__ __launch_bounds__(64) void mas_invert_kernel(const double* __restrict__ hess, TO* inv, int64_t C)
{
    constexpr int D = 3 * kBank;
    __shared__ double prow[2][2 * D];  // pivot row, stored twice: entry (p + k) needs no modulo
    const int         lane = threadIdx.x;
    const int64_t     k    = blockIdx.x;
    const bool        live = lane < D;
    const int         col  = live ? lane : 0;
    const int         nc = col / 3, a = col % 3;
    const double*     M  = hess + k * kCm;
    double            c[D];
#pragma unroll
    for(int row = 0; row < D; ++row)
    {
        const int nr = row / 3;
        double    v  = (nc >= nr) ? M[9 * sym_index(nr, nc) + (row % 3) + 3 * a]
                                  : M[9 * sym_index(nc, nr) + a + 3 * (row % 3)];
        if(row == col && v == 0.0)
            v = 1.0;
        c[row] = v;
    }
#pragma unroll 1
    for(int p = 0; p < D; ++p)
    {
        const double f = c[0];  // A(p, col)
        if(live)
        {
            prow[p & 1][lane]     = f;
            prow[p & 1][lane + D] = f;
        }
        __syncthreads();
        const double* sr    = &prow[p & 1][p];  // sr[q] = A(p, (p + q) mod 48)
        const double  inv_k = 1.0 / sr[0];
        const bool    piv   = (col == p);
        const double  s     = piv ? 0.0 : f * inv_k;
        const double  scale = piv ? inv_k : 1.0;
#pragma unroll
        for(int q = 1; q < D; ++q)
            c[q] = fma(-sr[q], s, c[q]) * scale;
        const double nv = piv ? -inv_k : s;
#pragma unroll
        for(int q = 0; q < D - 1; ++q)
            c[q] = c[q + 1];
        c[D - 1] = nv;
    }
    if(!live)
        return;
    TO* O = inv + k * kCm;
#pragma unroll
    for(int row = 0; row < D; ++row)
    {
        const int    nr = row / 3, b = row % 3;
        const double v  = -c[row];
        if(nr < nc)
            O[inv_at(sym_index(nr, nc), b + 3 * a)] = (TO)v;
        else if(nr == nc)
        {
            const int si = sym_index(nc, nc);
            if(b <= a)
                O[inv_at(si, b + 3 * a)] = (TO)v;
            if(b < a)
                O[inv_at(si, a + 3 * b)] = (TO)v;
        }
    }
}

// The same inverse on FP64 matrix cores: one wave per cluster, the 48x48
// matrix resident in MFMA accumulators (3x3 tiles of 16x16; tile (I, J)
// entry v of lane 16g + c holds S(16I + 4v + g, 16J + c)), swept in twelve
// pivot panels of 4 rows.  For panel K, with P = S_KK = L L^T (Cholesky):
//   W   = L^-1 S_K,:  (its K columns zero)
//   S  -= W^T W                    nine v_mfma_f64_16x16x4, one per tile
//   S_K,: = L^-T W,  S_:,K = its transpose,  S_KK = -L^-T L^-1
// which is the symmetric sweep S_ij -= S_iK P^-1 S_Kj, S_Kj <- P^-1 S_Kj,
// S_KK <- -P^-1 with the Schur update in its Cholesky form: the explicit
// inverse of a pivot block (S_iK P^-1 S_Kj) loses up to all digits on
// clusters with stiff contact couplings (nearly rank-one pivot blocks),
// the Cholesky form keeps the column sweep's accuracy.  After the twelve
// panels S = -A^-1.  The cluster's blocks come in and its inverse goes out
// through a 9.8 KB LDS copy (coalesced global accesses; scattered ones held
// 36 addresses per lane).  Both MFMA operands of tile (I, J) are lane values of W
// (A = W_I^T: lane 16g + c holds W(g, 16I + c); B = W_J: W(g, 16J + c)), so
// a lane computes three values of W per panel; the panel's rows, W and
// L^-T W pass between lanes through 4.5 KB of LDS.  The 4x4 Cholesky
// factor and its inverse run redundantly on every lane.  No pivoting, as in
// the reference's sweep: the cluster matrices are SPD and a zero diagonal (a
// padding slot) is set to 1.
//
// v_mfma_f64_16x16x4 lane layout (lane l, g = l / 16, c = l % 16):
//   A operand A[c][g], B operand B[g][c], accumulator v holds D[4 v + g][c]
// (checked on gfx950 with distinct, unsymmetric operands).
typedef double mfma_f64x4 __attribute__((ext_vector_type(4)));

template <typename TO>
__global__ __launch_bounds__(64, 2) void mas_invert_mfma_kernel(const double* __restrict__ hess, TO* inv, int64_t C)
{
    constexpr int D = 3 * kBank;
    __shared__ double Ms[kCm];     // the cluster's upper blocks in, its inverse's planes out
    __shared__ double rowb[4][D];  // the panel's rows of S
    __shared__ double wb[4][D];    // W
    __shared__ double xb[4][D];    // L^-T W
    const int         lane = threadIdx.x;
    const int         g = lane >> 4, c = lane & 15;
    const int64_t     k = blockIdx.x;
    const double*     M = hess + k * kCm;
    // the upper blocks (9 column-major scalars each), staged by coalesced
    // loads, into both triangles
    for(int e = lane; e < kCm; e += kWave)
        Ms[e] = M[e];
    __syncthreads();
    mfma_f64x4 S[3][3];
#pragma unroll
    for(int I = 0; I < 3; ++I)
#pragma unroll
        for(int J = 0; J < 3; ++J)
#pragma unroll
            for(int v = 0; v < 4; ++v)
            {
                const int row = 16 * I + 4 * v + g, col = 16 * J + c;
                const int a = row / 3, ra = row % 3, b = col / 3, cb = col % 3;
                double val = a <= b ? Ms[9 * sym_index(a, b) + ra + 3 * cb] : Ms[9 * sym_index(b, a) + cb + 3 * ra];
                if(row == col && val == 0.0)
                    val = 1.0;
                S[I][J][v] = val;
            }
#pragma unroll
    for(int J = 0; J < 3; ++J)
        rowb[g][16 * J + c] = S[0][J][0];
    __syncthreads();
#pragma unroll
    for(int kp = 0; kp < D / 4; ++kp)
    {
        const int k0 = 4 * kp, Kt = k0 / 16, ko = k0 % 16, vK = ko / 4;
        // P = L L^T and L^-1 (lower triangular), on every lane: one rsqrt per
        // pivot (rl = 1 / L_mm), no divisions
        double L[4][4] = {}, Li[4][4] = {}, rl[4];
#pragma unroll
        for(int m = 0; m < 4; ++m)
        {
#pragma unroll
            for(int n = 0; n < m; ++n)
            {
                double t = rowb[m][k0 + n];
#pragma unroll
                for(int q = 0; q < n; ++q)
                    t -= L[m][q] * L[n][q];
                L[m][n] = t * rl[n];
            }
            double t = rowb[m][k0 + m];
#pragma unroll
            for(int q = 0; q < m; ++q)
                t -= L[m][q] * L[m][q];
            rl[m]   = rsqrt(t);
            L[m][m] = t * rl[m];
        }
#pragma unroll
        for(int m = 0; m < 4; ++m)
        {
            Li[m][m] = rl[m];
#pragma unroll
            for(int n = 0; n < m; ++n)
            {
                double t = 0.0;
#pragma unroll
                for(int q = n; q < m; ++q)
                    t += L[m][q] * Li[q][n];
                Li[m][n] = -t * rl[m];
            }
        }
        // the lane's row and column of L^-1 (g is a lane value: selects, not
        // an indexed array, keep Li in registers)
        double lrow[4], lcol[4];
#pragma unroll
        for(int m = 0; m < 4; ++m)
        {
            lrow[m] = g == 0 ? Li[0][m] : g == 1 ? Li[1][m] : g == 2 ? Li[2][m] : Li[3][m];
            lcol[m] = g == 0 ? Li[m][0] : g == 1 ? Li[m][1] : g == 2 ? Li[m][2] : Li[m][3];
        }
        // W(g, 16T + c) = sum_m Li[g][m] S(k0 + m, 16T + c), zero in the K columns
        double w[3];
#pragma unroll
        for(int T = 0; T < 3; ++T)
        {
            double t = 0.0;
#pragma unroll
            for(int m = 0; m < 4; ++m)
                t += lrow[m] * rowb[m][16 * T + c];
            w[T]             = (T == Kt && c >= ko && c < ko + 4) ? 0.0 : t;
            wb[g][16 * T + c] = w[T];
        }
        __syncthreads();
        // X(g, 16T + c) = sum_{q >= g} Li[q][g] W(q, 16T + c)
        double x[3];
#pragma unroll
        for(int T = 0; T < 3; ++T)
        {
            double t = 0.0;
#pragma unroll
            for(int q = 0; q < 4; ++q)
                t += lcol[q] * wb[q][16 * T + c];
            x[T]             = t;
            xb[g][16 * T + c] = t;
        }
        // S -= W^T W
#pragma unroll
        for(int I = 0; I < 3; ++I)
#pragma unroll
            for(int J = 0; J < 3; ++J)
                S[I][J] = __builtin_amdgcn_mfma_f64_16x16x4f64(-w[I], w[J], S[I][J], 0, 0, 0);
        __syncthreads();
        // the pivot rows (lane row k0 + g), then the pivot columns (lanes of
        // columns k0 .. k0 + 3), then the pivot block
#pragma unroll
        for(int J = 0; J < 3; ++J)
            S[Kt][J][vK] = x[J];
        if(c >= ko && c < ko + 4)
        {
            const int m = c - ko;
#pragma unroll
            for(int I = 0; I < 3; ++I)
#pragma unroll
                for(int v = 0; v < 4; ++v)
                    S[I][Kt][v] = xb[m][16 * I + 4 * v + g];
            double t = 0.0;
#pragma unroll
            for(int q = 0; q < 4; ++q)
                t += lcol[q] * (m == 0 ? Li[q][0] : m == 1 ? Li[q][1] : m == 2 ? Li[q][2] : Li[q][3]);
            S[Kt][Kt][vK] = -t;
        }
        if(kp + 1 < D / 4)
        {
            const int k1 = k0 + 4, Kt1 = k1 / 16, vK1 = (k1 % 16) / 4;
#pragma unroll
            for(int J = 0; J < 3; ++J)
                rowb[g][16 * J + c] = S[Kt1][J][vK1];
        }
        __syncthreads();
        __builtin_amdgcn_sched_barrier(0);  // no work moves between panels (register pressure)
    }
    // A^-1 = -S, upper blocks as 9 planes of 136 (inv_at) into LDS, then
    // coalesced stores; a diagonal block takes its upper triangle for both
    // halves
#pragma unroll
    for(int I = 0; I < 3; ++I)
#pragma unroll
        for(int J = 0; J < 3; ++J)
#pragma unroll
            for(int v = 0; v < 4; ++v)
            {
                const int    row = 16 * I + 4 * v + g, col = 16 * J + c;
                const int    a = row / 3, ra = row % 3, b = col / 3, cb = col % 3;
                const double val = -S[I][J][v];
                if(a < b)
                    Ms[inv_at(sym_index(a, b), ra + 3 * cb)] = val;
                else if(a == b && ra <= cb)
                {
                    const int si = sym_index(a, a);
                    Ms[inv_at(si, ra + 3 * cb)] = val;
                    if(ra < cb)
                        Ms[inv_at(si, cb + 3 * ra)] = val;
                }
            }
    __syncthreads();
    TO* O = inv + k * kCm;
    for(int e = lane; e < kCm; e += kWave)
        O[e] = (TO)Ms[e];
}

__device__ __forceinline__ void inv3(const double m[9], double o[9])
{
    double c00 = m[4] * m[8] - m[7] * m[5];
    double c01 = m[7] * m[2] - m[1] * m[8];
    double c02 = m[1] * m[5] - m[4] * m[2];
    double id  = 1.0 / (m[0] * c00 + m[3] * c01 + m[6] * c02);
    o[0]       = c00 * id;
    o[1]       = c01 * id;
    o[2]       = c02 * id;
    o[3]       = (m[6] * m[5] - m[3] * m[8]) * id;
    o[4]       = (m[0] * m[8] - m[6] * m[2]) * id;
    o[5]       = (m[3] * m[2] - m[0] * m[5]) * id;
    o[6]       = (m[3] * m[7] - m[6] * m[4]) * id;
    o[7]       = (m[6] * m[1] - m[0] * m[7]) * id;
    o[8]       = (m[0] * m[4] - m[3] * m[1]) * id;
}

// 3x3 fallback for un-partitioned vertices (fem_mas_preconditioner.cu:24-45)
__global__ void mas_unpart_kernel(const int32_t* row_ptr,
                                  const int32_t* cols,
                                  const double*  vals,
                                  int32_t        blk_off,
                                  const int32_t* real_to_part,
                                  int32_t        V,
                                  double*        D)
{
    for(int32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < V; v += gridDim.x * blockDim.x)
    {
        if(real_to_part[v] >= 0)
            continue;
        int32_t row = v + blk_off, u = row_ptr[row];
        double  o[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        if(u < row_ptr[row + 1] && cols[u] == row)
        {
            double m[9];
            for(int q = 0; q < 9; ++q)
                m[q] = vals[upper_at(u, q)];
            inv3(m, o);
        }
        for(int q = 0; q < 9; ++q)
            D[9 * (int64_t)v + q] = o[q];
    }
}

// ------------------------------------------------------------------ apply kernels

// Z = C^-1 R of one cluster on one wave (schwarz_local_solve, 1603-1700).
// Each lane owns the upper blocks lane, lane + 64, lane + 128; block (a, b)
// gives M_ab R_b to row node a and, off the diagonal, M_ab^T R_a to row node
// b.  Every (row node, column node) pair gets exactly one contribution in
// sC[column node][3 row node + component]; a row's sum runs over the column
// nodes in ascending order, so the result does not depend on scheduling.
template <typename T>
struct ClusterBlocks
{
    T m[3][9];
};

template <typename T>
__device__ __forceinline__ void cluster_load(const T* __restrict__ M, int lane, ClusterBlocks<T>& cb)
{
#pragma unroll
    for(int t = 0; t < 3; ++t)
    {
        const int blk = lane + kWave * t;
        if(blk < kSym)
        {
#pragma unroll
            for(int q = 0; q < 9; ++q)
                cb.m[t][q] = M[inv_at(blk, q)];
        }
    }
}

template <typename T>
__device__ __forceinline__ void cluster_contrib(const ClusterBlocks<T>& cb,
                                                int                     lane,
                                                const T (*sR)[4],
                                                T (*sC)[3 * kBank])
{
#pragma unroll
    for(int t = 0; t < 3; ++t)
    {
        const int blk = lane + kWave * t;
        if(blk < kSym)
        {
            int a, b;
            sym_rc(blk, a, b);
            const T* m  = cb.m[t];
            const T  b0 = sR[b][0], b1 = sR[b][1], b2 = sR[b][2];
#pragma unroll
            for(int i = 0; i < 3; ++i)
                sC[b][3 * a + i] = m[i] * b0 + m[i + 3] * b1 + m[i + 6] * b2;
            if(a != b)
            {
                const T a0 = sR[a][0], a1 = sR[a][1], a2 = sR[a][2];
#pragma unroll
                for(int j = 0; j < 3; ++j)
                    sC[a][3 * b + j] = m[3 * j] * a0 + m[3 * j + 1] * a1 + m[3 * j + 2] * a2;
            }
        }
    }
}

template <typename T>
__device__ __forceinline__ T cluster_rowsum(const T (*sC)[3 * kBank], int row3)
{
    T acc = sC[0][row3];
#pragma unroll
    for(int j = 1; j < kBank; ++j)
        acc += sC[j][row3];
    return acc;
}

struct LevelSpan
{
    int32_t begin[kMaxLvl + 1];
    int32_t count[kMaxLvl + 1];
};

// Restriction R_{l+1} = sum of children.  Every lane loads its own node's
// value (coalesced); the lowest node of each in-bank sibling group (sib = bank
// mask | 0x10000 on the lowest) adds the group's values, taken from the 16
// lanes of its bank (a DPP row) by row broadcasts in ascending slot order, and stores the
// parent's sum: a parent's children all live in one bank, so there are no
// atomics and the sum is deterministic.  At level 0 the values are r gathered
// through slot_vec (the vector index of a slot's vertex, in the order the
// vectors are in).  Level starts are bank-aligned, so a bank never straddles
// two waves.
// lane Q of the lane's 16-lane bank (a DPP row): row_newbcast, a VALU move
// (no LDS round trip as with a shuffle).  Every lane of the row must be active.
template <int Q>
__device__ __forceinline__ float bank_lane(float v)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x150 + Q, 0xf, 0xf,
                                                                 false));
}
template <int Q>
__device__ __forceinline__ double bank_lane(double v)
{
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), 0x150 + Q, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), 0x150 + Q, 0xf, 0xf, false);
    return __hiloint2double(hi, lo);
}
// s = sum of the bank's lanes q with bit q of m set, in ascending q
template <typename T, int Q = 0>
__device__ __forceinline__ void bank_sums(T v0, T v1, T v2, int32_t m, T& s0, T& s1, T& s2)
{
    if constexpr(Q == 0)
        s0 = s1 = s2 = T(0);
    if constexpr(Q < kBank)
    {
        const T w0 = bank_lane<Q>(v0), w1 = bank_lane<Q>(v1), w2 = bank_lane<Q>(v2);
        if(m & (1 << Q))
        {
            s0 += w0;
            s1 += w1;
            s2 += w2;
        }
        bank_sums<T, Q + 1>(v0, v1, v2, m, s0, s1, s2);
    }
}

template <typename T>
__device__ __forceinline__ void restrict_sum_store(T              v0,
                                                   T              v1,
                                                   T              v2,
                                                   int32_t        m,
                                                   T*             Rb,
                                                   const int32_t* __restrict__ parent,
                                                   int32_t        slot)
{
    T s0, s1, s2;
    bank_sums<T>(v0, v1, v2, m, s0, s1, s2);
    if(m & 0x10000)
    {
        const int64_t pp = 3 * (int64_t)parent[slot];
        Rb[pp]           = s0;
        Rb[pp + 1]       = s1;
        Rb[pp + 2]       = s2;
    }
}

// the same sums, with the slot's parent already loaded (par); the lane's sums
// are returned in s0, s1, s2 (zero unless m holds siblings)
template <typename T>
__device__ __forceinline__ void restrict_sum_store_to(T v0, T v1, T v2, int32_t m, T* Rb, int32_t par, T& s0, T& s1,
                                                      T& s2)
{
    bank_sums<T>(v0, v1, v2, m, s0, s1, s2);
    if(m & 0x10000)
    {
        const int64_t pp = 3 * (int64_t)par;
        Rb[pp]           = s0;
        Rb[pp + 1]       = s1;
        Rb[pp + 2]       = s2;
    }
}

template <typename T, bool kLevel0>
__device__ __forceinline__ void restrict_bank(const double* __restrict__ r,
                                              const int32_t* __restrict__ slot_vec,
                                              T*             Rb,
                                              const int32_t* __restrict__ parent,
                                              const int32_t* __restrict__ sib,
                                              int32_t        slot,
                                              bool           live)
{
    T       v0 = 0, v1 = 0, v2 = 0;
    int32_t m  = 0;
    if(live)
    {
        m = sib[slot];
        if constexpr(kLevel0)
        {
            const int32_t v = slot_vec[slot];
            if(v >= 0)
            {
                v0 = (T)r[3 * (int64_t)v];
                v1 = (T)r[3 * (int64_t)v + 1];
                v2 = (T)r[3 * (int64_t)v + 2];
            }
        }
        else
        {
            v0 = Rb[3 * (int64_t)slot];
            v1 = Rb[3 * (int64_t)slot + 1];
            v2 = Rb[3 * (int64_t)slot + 2];
        }
    }
    restrict_sum_store<T>(v0, v1, v2, m, Rb, parent, slot);
}

template <typename T, bool kLevel0>
__global__ __launch_bounds__(256) void mas_restrict_level_kernel(const double* __restrict__ r,
                                                                const int32_t* __restrict__ slot_vec,
                                                                T*             Rb,
                                                                const int32_t* __restrict__ parent,
                                                                const int32_t* __restrict__ sib,
                                                                int32_t        begin,
                                                                int32_t        cnt,
                                                                const int32_t* conv)
{
    if(conv && *conv != 0)
        return;
    const int32_t x = blockIdx.x * blockDim.x + threadIdx.x;
    restrict_bank<T, kLevel0>(r, slot_vec, Rb, parent, sib, begin + x, x < cnt);
}

// fused_update_xr (linear_fused_pcg.cu:504-549) with the level-0 restriction
// folded in: one lane per partition slot (a whole number of workgroups),
// then one lane per unpartitioned vertex.  A slot's lane updates x and r of
// its vertex and hands the new r to the sibling sums of restrict_bank, so r
// is not read back by a separate restriction kernel.
//
// With kLevel1 (every level-1 bank's parts lie in one 256-slot range, checked
// at the build) the workgroup also holds the whole level-1 bank its slots
// restrict to, and the bank's sibling sums give level 2's R as well.
template <typename T, bool kLevel1>
__global__ __launch_bounds__(256) void mas_update_xr_restrict_kernel(PcgScalars*   sc,
                                                                    const double* pap_part,
                                                                    int           gp,
                                                                    double*       x,
                                                                    const double* __restrict__ p,
                                                                    double*       r,
                                                                    const double* __restrict__ Ap,
                                                                    const int32_t* __restrict__ slot_vec,
                                                                    T*            Rb,
                                                                    const int32_t* __restrict__ parent,
                                                                    const int32_t* __restrict__ sib,
                                                                    int32_t       nslots,
                                                                    int32_t       slot_blocks,
                                                                    const int32_t* __restrict__ unpart_vec,
                                                                    int32_t       n_unpart,
                                                                    int32_t       begin1,
                                                                    int32_t       count1,
                                                                    const int32_t* __restrict__ range_bank,
                                                                    double*       zero,
                                                                    int           nzero)
{
    __shared__ double lds[256 / kWave];
    __shared__ T      r1s[kBank][3];
    // the lane's share of the SpMV's pAp partials, its indices, then its
    // vertex's x, p, r and Ap are loaded before alpha's re-sum, so that the
    // operands' latency overlaps the sum; the converged flag is tested once
    // the index loads are out.  (Loads are unconditional, from clamped
    // in-bounds indices, so that no branch makes the compiler wait for them.)
    const int32_t      conv = sc->converged;
    const double       rz   = sc->rz_pending;
    PartialsAhead<256> ahead;
    ahead.load(pap_part, gp);
    const bool    slot_blk = (int32_t)blockIdx.x < slot_blocks;
    const int32_t slot     = blockIdx.x * 256 + threadIdx.x;
    const bool    live     = slot_blk && slot < nslots;
    int32_t       v = -1, m = 0, par = 0;
    if(slot_blk)
    {
        const int32_t sl = live ? slot : nslots - 1;
        const int32_t vv = slot_vec[sl], mm = sib[sl], pp = parent[sl];
        v                = live ? vv : -1;
        m                = live ? mm : 0;
        par              = live ? pp : 0;
    }
    else
    {
        const int32_t u  = (blockIdx.x - slot_blocks) * 256 + threadIdx.x;
        const int32_t vv = unpart_vec[u < n_unpart ? u : n_unpart - 1];
        v                = u < n_unpart ? vv : -1;
    }
    // kLevel1: the workgroup's level-1 bank (from the build) and, for lanes
    // below kBank, its node's sibling mask and parent
    int32_t bank1 = -1, m1 = 0, par1 = 0;
    if constexpr(kLevel1)
    {
        if(slot_blk)
        {
            bank1              = range_bank[blockIdx.x];
            const int32_t node = min(max(bank1, 0) * kBank + (int)(threadIdx.x % kBank), count1 - 1);
            m1                 = sib[begin1 + node];
            par1               = parent[begin1 + node];
        }
    }
    __builtin_amdgcn_sched_barrier(0);
    if(conv)
        return;
    if(blockIdx.x == 0)
        for(int k = threadIdx.x; k < nzero; k += 256)
            zero[k] = 0.0;
    const int64_t i = 3 * (int64_t)(v < 0 ? 0 : v);
    double        xv[3], pv[3], rv[3], av[3];
#pragma unroll
    for(int c = 0; c < 3; ++c)
    {
        rv[c] = r[i + c];
        av[c] = Ap[i + c];
        pv[c] = p[i + c];
        xv[c] = x[i + c];
    }
    if(v < 0)
    {
#pragma unroll
        for(int c = 0; c < 3; ++c)
            rv[c] = 0.0;  // a dead lane adds nothing to the sibling sums
    }
    if constexpr(kLevel1)
    {
        // ordered before the writes below by the barriers of alpha's sum
        if(slot_blk && threadIdx.x < kBank)
            r1s[threadIdx.x][0] = r1s[threadIdx.x][1] = r1s[threadIdx.x][2] = T(0);
    }
    const double alpha = pcg_alpha_of<256>(sc, ahead.sum(pap_part, gp, lds), rz);
    if(v >= 0)
    {
#pragma unroll
        for(int c = 0; c < 3; ++c)
        {
            rv[c]    = rv[c] - alpha * av[c];
            x[i + c] = xv[c] + alpha * pv[c];
            r[i + c] = rv[c];
        }
    }
    if(!slot_blk)
        return;
    const T v0 = (T)rv[0], v1 = (T)rv[1], v2 = (T)rv[2];
    T       s0, s1, s2;
    restrict_sum_store_to<T>(v0, v1, v2, m, Rb, par, s0, s1, s2);
    if constexpr(kLevel1)
    {
        // the level-1 sums, also into LDS by the node's place in its bank
        if(m & 0x10000)
        {
            // every node of the workgroup lies in bank1 (checked at the build)
            const int32_t node   = par - begin1;
            r1s[node % kBank][0] = s0;
            r1s[node % kBank][1] = s1;
            r1s[node % kBank][2] = s2;
        }
        __syncthreads();
        if(threadIdx.x < kWave && bank1 >= 0)
        {
            const int     q    = threadIdx.x;
            const int32_t node = bank1 * kBank + q;
            const bool    in   = q < kBank && node < count1;
            T t0, t1, t2;
            restrict_sum_store_to<T>(in ? r1s[q][0] : T(0), in ? r1s[q][1] : T(0), in ? r1s[q][2] : T(0),
                                     in ? m1 : 0, Rb, par1, t0, t1, t2);
        }
    }
}

// one wave per coarse cluster: Z = C^-1 R on the restricted residual.  The
// restriction kernels stop at the base level b (the first level small enough);
// clusters above it take each node's R as the sum of its level-b descendants
// (desc_ptr / desc_idx, ascending).  A cluster with m live nodes gives each
// node lpn = 64 / m' lanes (m' = m rounded up to a power of two, from the
// host, desc_lpn); each lane sums a strided share, and the lpn shares are
// combined by a fixed xor tree.
template <typename T>
__global__ __launch_bounds__(64) void mas_apply_coarse_kernel(const T* __restrict__ inv,
                                                             T*             Rb,
                                                             T*             Zb,
                                                             int64_t        c_first,
                                                             int64_t        c_desc,
                                                             const int32_t* __restrict__ desc_ptr,
                                                             const int32_t* __restrict__ desc_idx,
                                                             const int32_t* __restrict__ desc_lpn,
                                                             const int32_t* conv)
{
    if(conv && *conv != 0)
        return;
    __shared__ T     sR[kBank][4];
    __shared__ T     sC[kBank][3 * kBank];
    const int        lane = threadIdx.x;
    const int64_t    k    = c_first + blockIdx.x;
    ClusterBlocks<T> cb;
    cluster_load<T>(inv + k * kCm, lane, cb);
    if(k < c_desc)
    {
        if(lane < 3 * kBank)
            sR[lane / 3][lane % 3] = Rb[3 * k * kBank + lane];
    }
    else
    {
        const int lpn  = desc_lpn[k - c_desc];  // 4, 8, 16, 32 or 64 (uniform)
        const int node = lane / lpn, part = lane % lpn;
        T         s0 = 0, s1 = 0, s2 = 0;
        if(node < kBank)
        {
            const int64_t q  = (k - c_desc) * kBank + node;
            const int32_t d0 = desc_ptr[q], d1 = desc_ptr[q + 1];
#pragma unroll 4
            for(int32_t d = d0 + part; d < d1; d += lpn)
            {
                const int64_t y = 3 * (int64_t)desc_idx[d];
                s0 += Rb[y];
                s1 += Rb[y + 1];
                s2 += Rb[y + 2];
            }
        }
        // the node's lpn lanes (an aligned group; lpn >= 4, uniform)
        s0 = group_sum(s0, lpn);
        s1 = group_sum(s1, lpn);
        s2 = group_sum(s2, lpn);
        if(lane < kBank)
        {
            sR[lane][0] = 0;
            sR[lane][1] = 0;
            sR[lane][2] = 0;
        }
        __syncthreads();
        if(part == 0 && node < kBank)
        {
            sR[node][0] = s0;
            sR[node][1] = s1;
            sR[node][2] = s2;
        }
    }
    __syncthreads();
    cluster_contrib<T>(cb, lane, sR, sC);
    __syncthreads();
    if(lane < 3 * kBank)
        Zb[3 * k * kBank + lane] = cluster_rowsum<T>(sC, lane);
}

// the PCG's r.z and r.r of one wave (each summed over the wave in a fixed
// order) added into dot slot blk % kDotSlots (pairs) with two atomics: the
// slots' sums are in arrival order, as the SpMV's in-window sums are
constexpr int kDotSlots = 256;
__device__ __forceinline__ void wave_dot_to_slots(double s0, double s1, double* slots, int64_t blk)
{
    s0 = wave_total(s0);
    s1 = wave_total(s1);
    if((threadIdx.x % kWave) == 0)
    {
        double* d = slots + 2 * (blk % kDotSlots);
        atomicAdd(d, s0);
        atomicAdd(d + 1, s1);
    }
}

// one wave per fine cluster: gather r, local solve, and the prolongation of
// every level (injection): z_i = Z0_i + sum_{l >= 1} Z_l[ancestor_l(i)].
// Blocks past the fine clusters take the un-partitioned vertices (3x3 inverse
// diagonal block, fem_mas_preconditioner.cu:24-45).  With dot_slots the
// wave also adds its r.z and r.r into them (wave_dot_to_slots).
template <typename T>
__global__ __launch_bounds__(64) void mas_apply_fine_kernel(const double* __restrict__ r,
                                                           double*        z,
                                                           const int32_t* __restrict__ slot_vec,
                                                           const T* __restrict__ inv,
                                                           const T* __restrict__ Zb,
                                                           const int32_t* __restrict__ anc_slot,
                                                           int32_t        N,
                                                           int            Lact,
                                                           int64_t        C0,
                                                           const int32_t* __restrict__ unpart,
                                                           const int32_t* __restrict__ unpart_vec,
                                                           int32_t        n_unpart,
                                                           const double* __restrict__ Dun,
                                                           const int32_t* conv,
                                                           const int32_t* __restrict__ on,
                                                           const int32_t* __restrict__ lvl_of,
                                                           double*        dot_slots)
{
    if(conv && *conv != 0)
        return;
    // a cluster whose systems have all stopped keeps its z (batched solves)
    if(on && (int64_t)blockIdx.x < C0 && on[blockIdx.x] == 0)
        return;
    const int lane = threadIdx.x;
    if((int64_t)blockIdx.x >= C0)
    {
        const int64_t u = ((int64_t)blockIdx.x - C0) * kWave + lane;
        double s0 = 0.0, s1 = 0.0;
        if(u < n_unpart)
        {
            const double* D  = Dun + 9 * (int64_t)unpart[u];
            const int64_t q  = 3 * (int64_t)unpart_vec[u];
            const double  r0 = r[q], r1 = r[q + 1], r2 = r[q + 2];
            for(int a = 0; a < 3; ++a)
            {
                const double za = D[a] * r0 + D[a + 3] * r1 + D[a + 6] * r2;
                z[q + a]        = za;
                s0 += (a == 0 ? r0 : a == 1 ? r1 : r2) * za;
            }
            s1 = r0 * r0 + r1 * r1 + r2 * r2;
        }
        if(dot_slots)
            wave_dot_to_slots(s0, s1, dot_slots, blockIdx.x);
        return;
    }
    __shared__ T  sR[kBank][4];
    __shared__ T  sC[kBank][3 * kBank];
    const int64_t k = blockIdx.x;
    // the cluster's own level count (batched systems with their own counts)
    const int La = lvl_of ? lvl_of[k] : Lact;
    // lane 3 i + c (< 48) owns component c of node i: its vertex and its
    // ancestors are loaded first, then the inverse stream is issued, then the
    // gathers of r and of the coarse Z, so the gathers overlap the stream
    const bool    own  = lane < 3 * kBank;
    const int     i    = lane / 3, c = lane % 3;
    const int64_t slot = k * kBank + i;
    int32_t       v    = -1;
    int32_t       anc[kMaxLvl - 1];
    if(own)
    {
        v = slot_vec[slot];
#pragma unroll
        for(int l = 1; l < kMaxLvl; ++l)
            anc[l - 1] = l < La ? anc_slot[(int64_t)(l - 1) * N + slot] : 0;
    }
    ClusterBlocks<T> cb;
    cluster_load<T>(inv + k * kCm, lane, cb);
    T      zl[kMaxLvl - 1];
    double rv = 0.0;
    if(own)
    {
        rv       = v >= 0 ? r[3 * (int64_t)v + c] : 0.0;
        sR[i][c] = (T)rv;
#pragma unroll
        for(int l = 1; l < kMaxLvl; ++l)
            zl[l - 1] = (v >= 0 && l < La) ? Zb[3 * (int64_t)anc[l - 1] + c] : T(0);
    }
    __syncthreads();
    cluster_contrib<T>(cb, lane, sR, sC);
    __syncthreads();
    double s0 = 0.0;
    if(own && v >= 0)
    {
        T acc = cluster_rowsum<T>(sC, lane);
#pragma unroll
        for(int l = 1; l < kMaxLvl; ++l)
            if(l < La)
                acc += zl[l - 1];
        const double zv       = (double)acc;
        z[3 * (int64_t)v + c] = zv;
        s0                    = rv * zv;
    }
    if(dot_slots)
        wave_dot_to_slots(s0, rv * rv, dot_slots, blockIdx.x);
}

// The fine apply of a MAS preconditioner that covers the whole vector, fused
// with the PCG's r.z and r.r partials (one pair per workgroup, summed in a
// fixed order by the p update): four waves per workgroup, each on its own
// cluster (or 64 unpartitioned vertices) per round.  Every vertex's z is
// written by exactly one wave, so the partials cover the vector once.
template <typename T>
__global__ __launch_bounds__(256) void mas_apply_fine_dot_kernel(const double* __restrict__ r,
                                                                double*        z,
                                                                const int32_t* __restrict__ slot_vec,
                                                                const T* __restrict__ inv,
                                                                const T* __restrict__ Zb,
                                                                const int32_t* __restrict__ anc_slot,
                                                                int32_t        N,
                                                                int            Lact,
                                                                int64_t        C0,
                                                                const int32_t* __restrict__ unpart,
                                                                const int32_t* __restrict__ unpart_vec,
                                                                int32_t        n_unpart,
                                                                const double* __restrict__ Dun,
                                                                const int32_t* conv,
                                                                double*        part,
                                                                const int32_t* __restrict__ lvl_of)
{
    if(conv && *conv != 0)
        return;
    constexpr int kWaves = 4;
    __shared__ T      sR[kWaves][kBank][4];
    __shared__ T      sC[kWaves][kBank][3 * kBank];
    __shared__ double lds[kWaves];
    const int         w = threadIdx.x / kWave, lane = threadIdx.x % kWave;
    const int64_t     nblk = C0 + ((int64_t)n_unpart + kWave - 1) / kWave;
    const bool        own  = lane < 3 * kBank;
    const int         i = lane / 3, c = lane % 3;
    double            s0 = 0.0, s1 = 0.0;
    for(int64_t base = (int64_t)blockIdx.x * kWaves; base < nblk; base += (int64_t)gridDim.x * kWaves)
    {
        const int64_t k    = base + w;
        const bool    clus = k < C0;
        if(!clus && k < nblk)
        {
            const int64_t u = (k - C0) * kWave + lane;
            if(u < n_unpart)
            {
                const double* D  = Dun + 9 * (int64_t)unpart[u];
                const int64_t q  = 3 * (int64_t)unpart_vec[u];
                const double  r0 = r[q], r1 = r[q + 1], r2 = r[q + 2];
                for(int a = 0; a < 3; ++a)
                {
                    const double za = D[a] * r0 + D[a + 3] * r1 + D[a + 6] * r2;
                    z[q + a]        = za;
                    s0 += (a == 0 ? r0 : a == 1 ? r1 : r2) * za;
                }
                s1 += r0 * r0 + r1 * r1 + r2 * r2;
            }
        }
        int32_t v  = -1;
        double  rv = 0.0;
        T       zl[kMaxLvl - 1];
        const int La = (clus && lvl_of) ? lvl_of[k] : Lact;  // the cluster's own level count
        if(clus)
        {
            const int64_t slot = k * kBank + i;
            int32_t       anc[kMaxLvl - 1];
            if(own)
            {
                v = slot_vec[slot];
#pragma unroll
                for(int l = 1; l < kMaxLvl; ++l)
                    anc[l - 1] = l < La ? anc_slot[(int64_t)(l - 1) * N + slot] : 0;
            }
            ClusterBlocks<T> cb;
            cluster_load<T>(inv + k * kCm, lane, cb);
            if(own)
            {
                rv          = v >= 0 ? r[3 * (int64_t)v + c] : 0.0;
                sR[w][i][c] = (T)rv;
#pragma unroll
                for(int l = 1; l < kMaxLvl; ++l)
                    zl[l - 1] = (v >= 0 && l < La) ? Zb[3 * (int64_t)anc[l - 1] + c] : T(0);
            }
            __syncthreads();
            cluster_contrib<T>(cb, lane, sR[w], sC[w]);
        }
        else
            __syncthreads();
        __syncthreads();
        if(clus && own && v >= 0)
        {
            T acc = cluster_rowsum<T>(sC[w], lane);
#pragma unroll
            for(int l = 1; l < kMaxLvl; ++l)
                if(l < La)
                    acc += zl[l - 1];
            const double zv        = (double)acc;
            z[3 * (int64_t)v + c] = zv;
            s0 += rv * zv;
            s1 += rv * rv;
        }
        __syncthreads();  // sR and sC are rewritten next round
    }
    s0 = block_sum<256>(s0, lds);
    s1 = block_sum<256>(s1, lds);
    if(threadIdx.x == 0)
    {
        part[2 * blockIdx.x]     = s0;
        part[2 * blockIdx.x + 1] = s1;
    }
}

int grid_for(int64_t n, int threads = 256)
{
    return std::max(1, std::min(ceil_div(n, threads), device_cu_count() * 8));
}

template <typename T>
DBuf<T>& upload(DBuf<T>& d, const std::vector<T>& h)
{
    d.resize(std::max<size_t>(1, h.size()));
    if(!h.empty())
        MI_HIP(hipMemcpy(d.data(), h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice));
    return d;
}

class MasPreconditioner final : public LocalPreconditioner
{
  public:
    Hierarchy h;
    int       fp32 = 0, active_levels = 0, invert = MI_MAS_INVERT_MFMA;
    int64_t   C = 0, C0 = 0;  // clusters, fine clusters
    // device tables
    DBuf<int32_t> d_part_to_real, d_real_to_part, d_anc, d_anc_slot, d_parent, d_sib, d_unpart;
    DBuf<int32_t> d_range_bank;  // l1_local_: the level-1 bank of every 256-slot range (-1: none)
    // the vector index of every slot and unpartitioned vertex when the
    // vectors are in the SpMV's row order (the part order: rank_)
    DBuf<int32_t> d_part_to_pos, d_unpart_pos;
    bool          row_order_ = false;
    DBuf<int32_t> d_desc_ptr, d_desc_idx, d_desc_lpn;
    int32_t       n_unpart  = 0;
    int           base_lvl  = 1;  // restriction kernels produce R up to this level
    int64_t       c_desc    = 0;  // first cluster above the base level
    LevelSpan     span{};
    int64_t       c_first = 0;  // first coarse cluster
    DBuf<double>  d_hess, d_inv64, d_Dun;
    DBuf<float>   d_inv32;
    DBuf<double>  d_R64, d_Z64;
    DBuf<float>   d_R32, d_Z32;
    std::vector<std::pair<int64_t, int64_t>> level_clusters;  // [first, count] per level
    std::vector<int32_t> rank_;        // part-order rank of every vertex
    bool           r0_fused_ = false;  // the PCG's x, r update wrote level 1's R (update_xr)
    bool           r1_fused_ = false;  // ... and level 2's
    bool           l1_local_ = false;  // every level-1 bank's parts lie in one 256-slot range
    DBuf<int32_t>  d_fine_on;          // per fine cluster: 1 while it holds a running system
    const int32_t* fine_on = nullptr;  // d_fine_on during a batched solve, else nullptr
    DBuf<int32_t>  d_fine_lvl;         // per fine cluster: its system's level count (set_system_levels)
    const int32_t* fine_lvl = nullptr;  // d_fine_lvl while per-system counts are set, else nullptr
    int            lvl_max  = 0;        // the largest per-system count (the levels restricted and solved)

    explicit MasPreconditioner(const mi_precond_desc& d)
    {
        MI_REQUIRE(d.mas_build == MI_MAS_BUILD_HOST || d.mas_build == MI_MAS_BUILD_DEVICE, MI_ERR_ARG,
                   "MAS: unknown build location");
        const bool on_device = d.mas_build == MI_MAS_BUILD_DEVICE;
        MI_REQUIRE(d.part_ids_host != nullptr || (on_device && d.part_ids_dev != nullptr), MI_ERR_ARG,
                   "MAS: partition ids are required");
        MI_REQUIRE(d.vert_num == d.dof_count / 3, MI_ERR_ARG, "MAS: vert_num must equal dof_count / 3");
        MI_REQUIRE(d.mas_precision == MI_FP64 || d.mas_precision == MI_FP32, MI_ERR_ARG, "MAS: unknown precision");
        dof_offset    = d.dof_offset;
        dof_count     = d.dof_count;
        fp32          = d.mas_precision == MI_FP32;
        active_levels = d.mas_active_levels;
        invert        = d.mas_invert;
        MI_REQUIRE(invert == MI_MAS_INVERT_MFMA || invert == MI_MAS_INVERT_SWEEP, MI_ERR_ARG,
                   "MAS: unknown inverse kernel");
        h             = on_device ? mas_build_hierarchy_device(d) : build_hierarchy(d);
        // the part order of the vertices (the SpMV's window layout)
        rank_.assign((size_t)h.V, -1);
        {
            int32_t next = 0;
            for(int32_t s = 0; s < h.N; ++s)
                if(h.part_to_real[s] >= 0)
                    rank_[h.part_to_real[s]] = next++;
            for(int32_t v = 0; v < h.V; ++v)
                if(rank_[v] < 0)
                    rank_[v] = next++;
        }
        const int L   = h.L;
        C             = h.total / kBank;
        C0            = h.N / kBank;
        // per-vertex ancestors at levels 1..L-1 (global slots)
        std::vector<int32_t> anc((size_t)std::max(1, L - 1) * h.V, -1);
        for(int32_t v = 0; v < h.V; ++v)
            for(int l = 1; l < L; ++l)
                anc[(size_t)(l - 1) * h.V + v] = h.table[(size_t)kMaxLvl * v + (l - 1)];
        // parent of every slot (level 0 slots through their vertex)
        std::vector<int32_t> parent((size_t)std::max(1, h.total), -1);
        for(int32_t s = 0; s < h.N; ++s)
            if(h.part_to_real[s] >= 0)
                parent[s] = h.going_next[h.part_to_real[s]];
        for(int l = 1; l < L - 1; ++l)
            for(int32_t i = 0; i < h.level_count[l]; ++i)
                parent[h.level_begin[l] + i] = h.going_next[h.level_begin[l] + i];
        // clusters of every level
        level_clusters.clear();
        level_clusters.push_back({0, C0});
        for(int l = 1; l < L; ++l)
            level_clusters.push_back({h.level_begin[l] / kBank, bank_align(h.level_count[l]) / kBank});
        // sibling groups: bank mask of the nodes sharing a parent, bit 16 on
        // the lowest of them (level 0: the partition slots that hold a vertex)
        std::vector<int32_t> sib((size_t)std::max(1, h.total), 0);
        for(int l = 0; l + 1 < L; ++l)
        {
            const int32_t lb = l == 0 ? 0 : h.level_begin[l], lc = l == 0 ? h.N : h.level_count[l];
            for(int32_t i = 0; i < lc; ++i)
            {
                const int32_t slot = lb + i, b0 = slot & ~(kBank - 1);
                if(parent[slot] < 0)
                    continue;
                int32_t mask = 0, low = -1;
                for(int q = 0; q < kBank; ++q)
                {
                    const int32_t y = b0 + q;
                    if(y - lb < lc && parent[y] == parent[slot])
                    {
                        mask |= 1 << q;
                        if(low < 0)
                            low = q;
                    }
                }
                sib[slot] = mask | ((slot - b0) == low ? 0x10000 : 0);
            }
        }
        for(int l = 0; l <= L && l <= kMaxLvl; ++l)
        {
            span.begin[l] = l == 0 ? 0 : h.level_begin[l];
            span.count[l] = l == 0 ? h.N : h.level_count[l];
        }
        // ancestors by partition slot (the apply reads them in slot order)
        std::vector<int32_t> anc_slot((size_t)std::max(1, (L - 1) * h.N), 0);
        for(int32_t slot = 0; slot < h.N; ++slot)
            if(h.part_to_real[slot] >= 0)
                for(int l = 1; l < L; ++l)
                    anc_slot[(size_t)(l - 1) * h.N + slot] = anc[(size_t)(l - 1) * h.V + h.part_to_real[slot]];
        std::vector<int32_t> unpart;
        for(int32_t v = 0; v < h.V; ++v)
            if(h.real_to_part[v] < 0)
                unpart.push_back(v);
        n_unpart = (int32_t)unpart.size();
        upload(d_anc_slot, anc_slot);
        upload(d_unpart, unpart);
        {
            std::vector<int32_t> part_pos((size_t)h.N, -1), unpart_pos(unpart.size());
            for(int32_t slot = 0; slot < h.N; ++slot)
                if(h.part_to_real[slot] >= 0)
                    part_pos[slot] = rank_[h.part_to_real[slot]];
            for(size_t u = 0; u < unpart.size(); ++u)
                unpart_pos[u] = rank_[unpart[u]];
            upload(d_part_to_pos, part_pos);
            upload(d_unpart_pos, unpart_pos);
        }
        // base level: the first level at most kTailLevel nodes (restriction
        // kernels stop there); level-b descendants of every node above it
        base_lvl = L - 1;
        for(int l = 1; l < L; ++l)
            if(h.level_count[l] <= kTailLevel)
            {
                base_lvl = l;
                break;
            }
        const int32_t hi_begin = base_lvl + 1 < L ? h.level_begin[base_lvl + 1] : h.total;
        c_desc                 = hi_begin / kBank;
        const int32_t        hi_slots = h.total - hi_begin;
        std::vector<int32_t> cnt((size_t)hi_slots + 1, 0), desc_ptr((size_t)hi_slots + 1, 0), desc_idx;
        auto                 walk = [&](auto&& visit) {
            for(int32_t i = 0; i < h.level_count[base_lvl]; ++i)
            {
                const int32_t sb = h.level_begin[base_lvl] + i;
                for(int32_t a = parent[sb]; a >= 0; a = parent[a])
                    visit(a - hi_begin, sb);
            }
        };
        walk([&](int32_t q, int32_t) { ++cnt[q]; });
        for(int32_t q = 0; q < hi_slots; ++q)
            desc_ptr[q + 1] = desc_ptr[q] + cnt[q];
        desc_idx.resize((size_t)desc_ptr[hi_slots]);
        std::vector<int32_t> fillp(desc_ptr.begin(), desc_ptr.end() - 1);
        walk([&](int32_t q, int32_t sb) { desc_idx[fillp[q]++] = sb; });
        // lanes per node for every cluster above the base level
        std::vector<int32_t> lpn((size_t)std::max<int64_t>(1, C - c_desc), 4);
        for(int64_t k = c_desc; k < C; ++k)
        {
            int live = 0;
            for(int i = 0; i < kBank; ++i)
            {
                const int64_t q = (k - c_desc) * kBank + i;
                if(desc_ptr[q + 1] > desc_ptr[q])
                    live = i + 1;
            }
            int m = 1;
            while(m < live)
                m <<= 1;
            lpn[k - c_desc] = std::max(4, kWave / m);
        }
        upload(d_desc_ptr, desc_ptr);
        upload(d_desc_idx, desc_idx);
        upload(d_desc_lpn, lpn);
        c_first = L > 1 ? h.level_begin[1] / kBank : C;
        upload(d_sib, sib);
        upload(d_part_to_real, h.part_to_real);
        upload(d_real_to_part, h.real_to_part);
        upload(d_anc, anc);
        upload(d_parent, parent);
        // can one workgroup of 256 partition slots restrict level 1 too? Every
        // level-1 bank's nodes must come from slots of one 256-slot range, and
        // a range must feed at most one bank
        l1_local_ = L > 2;
        if(l1_local_)
        {
            const int32_t ranges = (int32_t)((h.N + 255) / 256);
            std::vector<int32_t> bank_range((size_t)std::max(1, (h.level_count[1] + kBank - 1) / kBank), -1),
                range_bank((size_t)std::max(1, ranges), -1);
            for(int32_t slot = 0; slot < h.N && l1_local_; ++slot)
            {
                if(h.part_to_real[slot] < 0)
                    continue;
                const int32_t node = parent[slot] - h.level_begin[1], bank = node / kBank, range = slot / 256;
                if(bank_range[bank] < 0)
                    bank_range[bank] = range;
                if(range_bank[range] < 0)
                    range_bank[range] = bank;
                l1_local_ = bank_range[bank] == range && range_bank[range] == bank;
            }
            if(l1_local_)
                upload(d_range_bank, range_bank);
        }
        d_hess.resize((size_t)std::max<int64_t>(1, C * kCm));
        if(fp32)
        {
            d_inv32.resize((size_t)std::max<int64_t>(1, C * kCm));
            d_R32.resize((size_t)std::max(1, 3 * h.total));
            d_Z32.resize((size_t)std::max(1, 3 * h.total));
            MI_HIP(hipMemset(d_R32.data(), 0, d_R32.bytes()));
        }
        else
        {
            d_inv64.resize((size_t)std::max<int64_t>(1, C * kCm));
            d_R64.resize((size_t)std::max(1, 3 * h.total));
            d_Z64.resize((size_t)std::max(1, 3 * h.total));
            MI_HIP(hipMemset(d_R64.data(), 0, d_R64.bytes()));
        }
        d_Dun.resize((size_t)std::max(1, 9 * h.V));
        MI_HIP(hipMemset(d_Dun.data(), 0, d_Dun.bytes()));
    }

    void assemble(const BsrMatrix& A, hipStream_t s) override
    {
        const int32_t blk_off = (int32_t)(dof_offset / 3);
        MI_REQUIRE(blk_off + h.V <= A.B, MI_ERR_ARG, "MAS: subsystem range exceeds the matrix");
        if(C == 0)
            return;
        fill_f64_kernel<<<grid_for(C * kCm), 256, 0, s>>>(d_hess.data(), C * kCm);
        bool range_walk = false;
        if(A.U > 0)
        {
            // in the partition order when every level-1 bank is one 256-slot
            // range (its cluster then summed in LDS); MI_MAS_WALK=blocks keeps
            // the block-order walk
            static const bool by_blocks = [] {
                const char* e = std::getenv("MI_MAS_WALK");
                return e && std::string(e) == "blocks";
            }();
            range_walk = l1_local_ && !by_blocks;
            if(range_walk)
                mas_walk_range_kernel<<<ceil_div(h.N, 256), 256, 0, s>>>(
                    A.row_ptr.data(), A.cols.data(), A.vals.data(), blk_off, h.V, d_part_to_real.data(),
                    d_real_to_part.data(), d_anc.data(), h.L, d_range_bank.data(), c_first, h.N, d_parent.data(),
                    d_hess.data());
            else
                mas_walk_kernel<<<ceil_div(A.U, 256), 256, 0, s>>>(A.U,
                                                                   A.rows.data(),
                                                                   A.cols.data(),
                                                                   A.vals.data(),
                                                                   blk_off,
                                                                   h.V,
                                                                   d_real_to_part.data(),
                                                                   d_anc.data(),
                                                                   h.L,
                                                                   d_hess.data());
        }
        // (the partition-order walk has done level 0's push)
        for(int l = range_walk ? 1 : 0; l + 1 < h.L; ++l)
        {
            auto [first, cnt] = level_clusters[l];
            if(cnt > 0)
                mas_push_kernel<<<ceil_div(cnt, 4), 256, 0, s>>>(first, cnt, d_parent.data(), d_hess.data());
        }
        if(invert == MI_MAS_INVERT_SWEEP)
        {
            if(fp32)
                mas_invert_kernel<float><<<(unsigned)C, 64, 0, s>>>(d_hess.data(), d_inv32.data(), C);
            else
                mas_invert_kernel<double><<<(unsigned)C, 64, 0, s>>>(d_hess.data(), d_inv64.data(), C);
        }
        else if(fp32)
            mas_invert_mfma_kernel<float><<<(unsigned)C, 64, 0, s>>>(d_hess.data(), d_inv32.data(), C);
        else
            mas_invert_mfma_kernel<double><<<(unsigned)C, 64, 0, s>>>(d_hess.data(), d_inv64.data(), C);
        if(h.has_unpart)
            mas_unpart_kernel<<<grid_for(h.V), 256, 0, s>>>(A.row_ptr.data(), A.cols.data(), A.vals.data(),
                                                            blk_off, d_real_to_part.data(), h.V, d_Dun.data());
        MI_HIP(hipGetLastError());
    }

    // the fine kernel writes the PCG's dot partials when `part` is given and
    // returns their count (at most max_parts)
    // part: the PCG's r.z and r.r partials, written by the persistent fused
    // kernel (at most max_parts pairs, their count returned), or, with
    // slots, added into kDotSlots zeroed pairs with atomics by the plain one
    template <typename T>
    int apply_t(const double* r, double* z, const int32_t* conv, hipStream_t s, const T* inv, T* R, T* Z,
                double* part = nullptr, int max_parts = 0, bool slots = false)
    {
        const int L    = h.L;
        const int Lact = levels_run();
        const int32_t* pv = slot_vec();
        const int32_t* uv = unpart_vec();
        // one active level is the 16-node block Jacobi: no coarse level is read
        const bool r0_done = r0_fused_, r1_done = r1_fused_;
        r0_fused_ = r1_fused_ = false;
        for(int l = 0; Lact > 1 && l < base_lvl; ++l)
        {
            const int32_t cnt = span.count[l];
            if(cnt == 0 || (l == 0 && r0_done) || (l == 1 && r1_done))
                continue;
            if(l == 0)
                mas_restrict_level_kernel<T, true><<<ceil_div(cnt, 256), 256, 0, s>>>(
                    r, pv, R, d_parent.data(), d_sib.data(), span.begin[l], cnt, conv);
            else
                mas_restrict_level_kernel<T, false><<<ceil_div(cnt, 256), 256, 0, s>>>(
                    r, pv, R, d_parent.data(), d_sib.data(), span.begin[l], cnt, conv);
        }
        if(Lact > 1 && C > c_first)
            mas_apply_coarse_kernel<T><<<(unsigned)(C - c_first), 64, 0, s>>>(
                inv, R, Z, c_first, c_desc, d_desc_ptr.data(), d_desc_idx.data(), d_desc_lpn.data(), conv);
        const int64_t blocks = C0 + ceil_div(n_unpart, kWave);
        if(part && !slots)
        {
            const int g = (int)std::max<int64_t>(1, std::min<int64_t>(max_parts, ceil_div(blocks, (int64_t)4)));
            mas_apply_fine_dot_kernel<T><<<g, 256, 0, s>>>(r, z, pv, inv, Z, d_anc_slot.data(), h.N, Lact, C0,
                                                           d_unpart.data(), uv, n_unpart, d_Dun.data(), conv, part,
                                                           fine_lvl);
            return g;
        }
        if(blocks > 0)
            mas_apply_fine_kernel<T><<<(unsigned)blocks, 64, 0, s>>>(r, z, pv, inv, Z, d_anc_slot.data(), h.N,
                                                                     Lact, C0, d_unpart.data(), uv, n_unpart,
                                                                     d_Dun.data(), conv, fine_on, fine_lvl,
                                                                     slots ? part : nullptr);
        return slots ? kDotSlots : 0;
    }

    bool mas_export(int32_t* info,
                    int32_t* p2r,
                    int32_t* r2p,
                    int32_t* gn,
                    int32_t* ls,
                    int32_t* ct,
                    double*  hess,
                    double*  inv) override
    {
        MI_HIP(hipDeviceSynchronize());
        if(info)
        {
            int32_t v[6] = {h.V, h.N, h.L, h.total, h.level1_begin, (int32_t)C};
            std::memcpy(info, v, sizeof(v));
        }
        if(p2r)
            std::memcpy(p2r, h.part_to_real.data(), sizeof(int32_t) * h.N);
        if(r2p)
            std::memcpy(r2p, h.real_to_part.data(), sizeof(int32_t) * h.V);
        if(gn)
            std::memcpy(gn, h.going_next.data(), sizeof(int32_t) * h.going_next.size());
        if(ls)
            for(int l = 0; l <= h.L; ++l)
            {
                ls[2 * l]     = h.level_count[l];
                ls[2 * l + 1] = h.level_begin[l];
            }
        if(ct)
            std::memcpy(ct, h.table.data(), sizeof(int32_t) * h.table.size());
        if(hess && C > 0)
            MI_HIP(hipMemcpy(hess, d_hess.data(), sizeof(double) * C * kCm, hipMemcpyDeviceToHost));
        if(inv && C > 0)
        {
            // device planes -> 9 scalars per upper block, block after block
            std::vector<double> t((size_t)C * kCm);
            if(fp32)
            {
                std::vector<float> f((size_t)C * kCm);
                MI_HIP(hipMemcpy(f.data(), d_inv32.data(), sizeof(float) * f.size(), hipMemcpyDeviceToHost));
                std::copy(f.begin(), f.end(), t.begin());
            }
            else
                MI_HIP(hipMemcpy(t.data(), d_inv64.data(), sizeof(double) * t.size(), hipMemcpyDeviceToHost));
            for(int64_t k = 0; k < C; ++k)
                for(int blk = 0; blk < kSym; ++blk)
                    for(int q = 0; q < 9; ++q)
                        inv[k * kCm + 9 * blk + q] = t[k * kCm + inv_at(blk, q)];
        }
        return true;
    }

    bool set_active_levels(int levels) override
    {
        active_levels = levels;  // 0 or out of range: every level, as the reference clamps
        fine_lvl      = nullptr;  // one count for every system again
        return true;
    }

    // the level count the restriction and the coarse solves run to
    int levels_run() const
    {
        if(fine_lvl)
            return lvl_max;
        return (active_levels > 0 && active_levels <= h.L) ? active_levels : h.L;
    }

    // batched solves: system g (block rows [seg[g], seg[g+1]), global) uses
    // levels[g] levels (0 or out of range: every level); a fine cluster takes
    // its system's count (the largest of its systems' should a cluster hold
    // vertices of several); nseg = 0 returns to the one count
    bool set_system_levels(int32_t nseg, const int32_t* seg, const int32_t* levels) override
    {
        if(nseg <= 0)
        {
            fine_lvl = nullptr;
            return true;
        }
        const int32_t        off = (int32_t)(dof_offset / 3);
        std::vector<int32_t> lv((size_t)std::max<int64_t>(C0, 1), 1);
        int                  mx = 1;
        for(int64_t k = 0; k < C0; ++k)
            for(int i = 0; i < kBank; ++i)
            {
                const int32_t v = h.part_to_real[k * kBank + i];
                if(v < 0)
                    continue;
                const int32_t  row = off + v;
                const int32_t* it  = std::upper_bound(seg, seg + nseg + 1, row);
                MI_REQUIRE(it != seg && it != seg + nseg + 1, MI_ERR_ARG,
                           "mi_precond_mas_set_system_levels: a vertex outside the systems' rows");
                const int32_t g  = (int32_t)(it - seg) - 1;
                const int     lg = (levels[g] > 0 && levels[g] <= h.L) ? levels[g] : h.L;
                lv[k]            = std::max<int32_t>(lv[k], lg);
                mx               = std::max(mx, lg);
            }
        upload(d_fine_lvl, lv);
        MI_HIP(hipDeviceSynchronize());  // the host vector goes out of scope
        fine_lvl = d_fine_lvl.data();
        lvl_max  = mx;
        return true;
    }

    const std::vector<int32_t>* spmv_rank() const override { return &rank_; }

    void set_row_order(bool on) override { row_order_ = on; }
    // the vector index of every partition slot (-1: empty) and of every
    // unpartitioned vertex, in the order the vectors are in
    const int32_t* slot_vec() const { return row_order_ ? d_part_to_pos.data() : d_part_to_real.data(); }
    const int32_t* unpart_vec() const { return row_order_ ? d_unpart_pos.data() : d_unpart.data(); }

    // the PCG's x, r update with the level-0 restriction folded in, when the
    // apply restricts at all (more than one active level) and this MAS covers
    // the whole vector (the caller checks the range); MI_MAS_FUSE_XR=0 keeps
    // the separate kernels
    bool update_xr(PcgScalars* sc, const double* pap_part, int gp, double* x, const double* p, double* r,
                   const double* Ap, hipStream_t s, double* zero, int nzero) override
    {
        static const int on = [] {
            const char* e = std::getenv("MI_MAS_FUSE_XR");
            return e ? std::atoi(e) : 1;
        }();
        const int Lact = levels_run();
        if(!on || h.V == 0 || Lact <= 1 || base_lvl < 1 || span.count[0] == 0)
            return false;
        const int32_t slot_blocks = ceil_div(span.count[0], 256);
        const int32_t blocks      = slot_blocks + ceil_div(n_unpart, 256);
        // level 2's R too, when level 1 is restricted by a kernel of the apply
        // (base level above 1) and its banks lie in whole workgroups
        static const int l1 = [] {
            const char* e = std::getenv("MI_MAS_FUSE_R1");
            return e ? std::atoi(e) : 1;
        }();
        const bool lvl1 = l1 && l1_local_ && base_lvl >= 2 && Lact > 2;
        auto go = [&](auto* R, auto tag) {
            using T = decltype(tag);
            auto launch = [&](auto kernel) {
                kernel<<<blocks, 256, 0, s>>>(sc, pap_part, gp, x, p, r, Ap, slot_vec(), R, d_parent.data(),
                                              d_sib.data(), span.count[0], slot_blocks, unpart_vec(), n_unpart,
                                              span.begin[1], span.count[1], lvl1 ? d_range_bank.data() : nullptr,
                                              zero, zero ? nzero : 0);
            };
            if(lvl1)
                launch(mas_update_xr_restrict_kernel<T, true>);
            else
                launch(mas_update_xr_restrict_kernel<T, false>);
        };
        if(fp32)
            go(d_R32.data(), float{});
        else
            go(d_R64.data(), double{});
        MI_HIP(hipGetLastError());
        r0_fused_ = true;
        r1_fused_ = lvl1;
        return true;
    }

    // fine clusters of stopped systems are skipped; the restriction and the
    // coarse levels (a few percent of the apply) still run over everything
    void skip_stopped(const int32_t* seg, int32_t nseg, const BatchScalars* sc, hipStream_t s) override
    {
        if(!seg || C0 == 0)
        {
            fine_on = nullptr;
            return;
        }
        d_fine_on.resize((size_t)C0);
        batch_cluster_flags(d_part_to_real.data(), C0, kBank, (int32_t)(dof_offset / 3), seg, nseg, sc,
                            d_fine_on.data(), s);
        fine_on = d_fine_on.data();
    }

    void apply(const double* r, double* z, const int32_t* conv, hipStream_t s) override
    {
        if(h.V == 0)
            return;
        if(fp32)
            apply_t<float>(r, z, conv, s, d_inv32.data(), d_R32.data(), d_Z32.data());
        else
            apply_t<double>(r, z, conv, s, d_inv64.data(), d_R64.data(), d_Z64.data());
    }

    // The r.z and r.r partials from the fine apply: by default each wave of
    // the plain fine kernel (one cluster) adds them into 256 dot slots with
    // FP64 atomics, the slots cleared by the x, r update before it (FP64
    // solve 176 -> 167 us per iteration, FP32 119 -> 117 us, against the
    // separate dot kernel and the persistent fused kernel respectively).
    // MI_MAS_DOT_SLOTS=0 restores the per-block partials.
    bool use_slots() const
    {
        static const int env = [] {
            const char* e = std::getenv("MI_MAS_DOT_SLOTS");
            return e ? std::atoi(e) : 1;
        }();
        return env != 0;
    }
    int dot_slots() const override { return h.V > 0 && use_slots() ? 2 * kDotSlots : 0; }

    int apply_dot(const double* r, double* z, const int32_t* conv, double* part, int max_parts,
                  hipStream_t s) override
    {
        const bool slots = use_slots();
        if(h.V == 0 || max_parts < (slots ? kDotSlots : 1))
            return -1;
        if(!fp32)
            return apply_t<double>(r, z, conv, s, d_inv64.data(), d_R64.data(), d_Z64.data(), part, max_parts, slots);
        return apply_t<float>(r, z, conv, s, d_inv32.data(), d_R32.data(), d_Z32.data(), part, max_parts, slots);
    }
};
}  // namespace

std::unique_ptr<LocalPreconditioner> make_mas(const mi_precond_desc& d)
{
    MI_REQUIRE(d.dof_offset % 3 == 0 && d.dof_count % 3 == 0, MI_ERR_ARG, "MAS: DoF range must be 3-aligned");
    return std::make_unique<MasPreconditioner>(d);
}
}  // namespace mi
