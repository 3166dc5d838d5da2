// MAS hierarchy built by device kernels (gfx950): the same tables as the host
// build in mas.hip, bit for bit, for topologies where the host adjacency and
// level loops are the cost of a (re)initialisation.
//
//   adjacency   every simplex emits its directed vertex pairs as keys
//               a * V + b (invalid pairs: the sentinel V * V), one stable radix
//               sort, run-length heads -> CSR (nstart, nnum, nlist): sorted,
//               unique neighbours per vertex, as the std::set / sort-unique of
//               FEMMASPreconditioner::do_init (fem_mas_preconditioner.cu:163-203);
//   slots       vertices of a partition collected by atomics (at most 16),
//               sorted per partition by one thread, partitions placed at
//               bank-aligned offsets (fem_mas_preconditioner.cu:291-356):
//               part_to_real / real_to_part in ascending vertex order;
//   levels      build_connect_mask_L0, prepare_prefix_sum_L0, build_level1,
//               build_connect_mask_Lx, next_level_cluster, prefix_sum_Lx,
//               compute_next_level, aggregation_kernel
//               (mas_preconditioner_engine.cu:57-466): one thread per node or
//               per 16-node bank; the neighbour lists are compacted in place
//               exactly as the host build does, and the coarse connect masks
//               are OR-ed with atomics (order-free).
// One host read-back per level for its node count, as the reference's init.
#include <algorithm>

#include "mas.hpp"

namespace mi
{
namespace
{
constexpr int kT = 256;

int grid_n(int64_t n)
{
    return std::max(1, std::min(ceil_div(n, kT), device_cu_count() * 16));
}

#define MI_GRID_STRIDE(i, n) \
    for(int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < (n); i += (int64_t)gridDim.x * blockDim.x)

__device__ __forceinline__ uint32_t lt16(int lane)
{
    return (1u << lane) - 1u;
}

__device__ __forceinline__ uint32_t closure16(const uint32_t cache[kMasBank], int lane, uint32_t start)
{
    uint32_t conn = start, visited = 1u << lane;
    for(;;)
    {
        uint32_t todo = visited ^ conn;
        if(!todo)
            break;
        int nv = __ffs(todo) - 1;
        visited |= 1u << nv;
        conn |= cache[nv];
    }
    return conn;
}

template <typename T>
__global__ void fill_value_kernel(T* p, int64_t n, T v)
{
    MI_GRID_STRIDE(i, n)
    p[i] = v;
}

__device__ __forceinline__ uint64_t pair_key(int32_t a, int32_t b, int32_t V)
{
    const bool ok = a != b && a >= 0 && b >= 0 && a < V && b < V;
    return ok ? (uint64_t)a * (uint64_t)V + (uint64_t)b : (uint64_t)V * (uint64_t)V;
}

// tets: 12 keys each, triangles 6, edges 2, in that order
__global__ void adj_emit_kernel(const int32_t* __restrict__ tets,
                                int64_t nt,
                                const int32_t* __restrict__ tris,
                                int64_t ntr,
                                const int32_t* __restrict__ edges,
                                int64_t  ne,
                                int32_t  V,
                                uint64_t* keys)
{
    MI_GRID_STRIDE(e, nt + ntr + ne)
    {
        int32_t   v[4];
        int       nv;
        uint64_t* out;
        if(e < nt)
        {
            nv  = 4;
            out = keys + 12 * e;
            for(int i = 0; i < 4; ++i)
                v[i] = tets[4 * e + i];
        }
        else if(e < nt + ntr)
        {
            const int64_t t = e - nt;
            nv              = 3;
            out             = keys + 12 * nt + 6 * t;
            for(int i = 0; i < 3; ++i)
                v[i] = tris[3 * t + i];
        }
        else
        {
            const int64_t t = e - nt - ntr;
            nv              = 2;
            out             = keys + 12 * nt + 6 * ntr + 2 * t;
            v[0]            = edges[2 * t];
            v[1]            = edges[2 * t + 1];
        }
        int q = 0;
        for(int i = 0; i < nv; ++i)
            for(int j = i + 1; j < nv; ++j)
            {
                out[q++] = pair_key(v[i], v[j], V);
                out[q++] = pair_key(v[j], v[i], V);
            }
    }
}

__device__ __forceinline__ bool adj_head(const uint64_t* keys, int64_t i, uint64_t VV)
{
    const uint64_t k = keys[i];
    return k < VV && (i == 0 || keys[i - 1] != k);
}

// head flags (scanned into output positions) and unique neighbours per vertex
__global__ void adj_head_kernel(const uint64_t* __restrict__ keys, int64_t E, int32_t V, uint32_t* head, uint32_t* deg)
{
    const uint64_t VV = (uint64_t)V * (uint64_t)V;
    MI_GRID_STRIDE(i, E)
    {
        const bool h = adj_head(keys, i, VV);
        head[i]      = h ? 1u : 0u;
        if(h)
            atomicAdd(&deg[keys[i] / (uint64_t)V], 1u);
    }
}

__global__ void adj_write_kernel(const uint64_t* __restrict__ keys,
                                 int64_t E,
                                 int32_t V,
                                 const uint32_t* __restrict__ pos,
                                 uint32_t* nlist)
{
    const uint64_t VV = (uint64_t)V * (uint64_t)V;
    MI_GRID_STRIDE(i, E)
    if(adj_head(keys, i, VV))
        nlist[pos[i]] = (uint32_t)(keys[i] % (uint64_t)V);
}

// stat[0] = max(0, max part id), stat[1] = 1 if a vertex is unpartitioned
__global__ void part_stat_kernel(const int32_t* __restrict__ part, int32_t V, int32_t* stat)
{
    MI_GRID_STRIDE(v, V)
    {
        const int32_t p = part[v];
        if(p < 0)
            stat[1] = 1;
        else
            atomicMax(&stat[0], p);
    }
}

// members of every partition (unordered, at most 16); stat[2] = 1 on overflow
__global__ void part_collect_kernel(const int32_t* __restrict__ part,
                                    int32_t  V,
                                    uint32_t* psize,
                                    int32_t*  members,
                                    int32_t*  stat)
{
    MI_GRID_STRIDE(v, V)
    {
        const int32_t p = part[v];
        if(p < 0)
            continue;
        const uint32_t k = atomicAdd(&psize[p], 1u);
        if(k < (uint32_t)kMasBank)
            members[(int64_t)kMasBank * p + k] = (int32_t)v;
        else
            stat[2] = 1;
    }
}

__global__ void part_align_kernel(const uint32_t* __restrict__ psize, int32_t P, uint32_t* asz)
{
    MI_GRID_STRIDE(p, P)
    asz[p] = (uint32_t)((psize[p] + kMasBank - 1) / kMasBank * kMasBank);
}

// slots of partition p: its vertices in ascending order from poff[p]
__global__ void part_fill_kernel(const uint32_t* __restrict__ psize,
                                 const uint32_t* __restrict__ poff,
                                 const int32_t* __restrict__ members,
                                 int32_t  P,
                                 int32_t* p2r,
                                 int32_t* r2p)
{
    MI_GRID_STRIDE(p, P)
    {
        const int n = (int)min(psize[p], (uint32_t)kMasBank);
        int32_t   m[kMasBank];
        for(int k = 0; k < n; ++k)
        {
            int32_t x = members[(int64_t)kMasBank * p + k];
            int     j = k;
            while(j > 0 && m[j - 1] > x)
            {
                m[j] = m[j - 1];
                --j;
            }
            m[j] = x;
        }
        for(int k = 0; k < n; ++k)
        {
            const int32_t s = (int32_t)poff[p] + k;
            p2r[s]          = m[k];
            r2p[m[k]]       = s;
        }
    }
}

// build_connect_mask_L0 (mas_preconditioner_engine.cu:57-106)
__global__ void l0_connect_kernel(int32_t N,
                                  const int32_t* __restrict__ p2r,
                                  const int32_t* __restrict__ r2p,
                                  const uint32_t* __restrict__ nstart,
                                  uint32_t* nnum,
                                  uint32_t* nlist,
                                  uint32_t* fine)
{
    MI_GRID_STRIDE(tid, N)
    {
        const int32_t idx = p2r[tid];
        if(idx < 0)
            continue;
        const int32_t  bank = (int32_t)(tid / kMasBank);
        uint32_t       msk  = 1u << (tid % kMasBank), nk = 0;
        const uint32_t st = nstart[idx], cnt = nnum[idx];
        for(uint32_t q = 0; q < cnt; ++q)
        {
            const uint32_t w  = nlist[st + q];
            const int32_t  np = r2p[w];
            if(np < 0)
                continue;
            if(np / kMasBank == bank)
                msk |= 1u << (np % kMasBank);
            else
                nlist[st + nk++] = w;
        }
        nnum[idx] = nk;
        fine[idx] = msk;
    }
}

// prepare_prefix_sum_L0 (108-156): closure inside each bank, elected count
__global__ void l0_prefix_kernel(int32_t nbanks, const int32_t* __restrict__ p2r, uint32_t* fine, uint32_t* pref)
{
    MI_GRID_STRIDE(b, nbanks)
    {
        uint32_t cache[kMasBank];
        int32_t  idx[kMasBank];
        for(int l = 0; l < kMasBank; ++l)
        {
            idx[l]   = p2r[b * kMasBank + l];
            cache[l] = idx[l] >= 0 ? fine[idx[l]] : 0u;
        }
        uint32_t cnt = 0;
        for(int l = 0; l < kMasBank; ++l)
        {
            if(idx[l] < 0)
                continue;
            const uint32_t c = closure16(cache, l, cache[l]);
            fine[idx[l]]     = c;
            cnt += __popc(c & lt16(l)) == 0;
        }
        pref[b] = cnt;
    }
}

// build_level1 (158-212): elected-lane prefix -> level-1 index and going_next
__global__ void level1_kernel(int32_t nbanks,
                              const int32_t* __restrict__ p2r,
                              const uint32_t* __restrict__ fine,
                              const uint32_t* __restrict__ psum,
                              int32_t  level1_begin,
                              int32_t* coarse0,
                              int32_t* going_next)
{
    MI_GRID_STRIDE(b, nbanks)
    {
        int32_t  idx[kMasBank];
        uint32_t elected = 0;
        for(int l = 0; l < kMasBank; ++l)
        {
            idx[l] = p2r[b * kMasBank + l];
            if(idx[l] >= 0 && __popc(fine[idx[l]] & lt16(l)) == 0)
                elected |= 1u << l;
        }
        for(int l = 0; l < kMasBank; ++l)
        {
            if(idx[l] < 0)
                continue;
            const int     el     = __ffs(fine[idx[l]]) - 1;
            const int32_t prefix = __popc(elected & lt16(el)) + (int32_t)psum[b];
            coarse0[idx[l]]      = prefix;
            going_next[idx[l]]   = prefix + level1_begin;
        }
    }
}

// build_connect_mask_Lx (214-294)
__global__ void lx_connect_kernel(int32_t N,
                                  const int32_t* __restrict__ p2r,
                                  const int32_t* __restrict__ cprev,
                                  const uint32_t* __restrict__ nstart,
                                  uint32_t* nnum,
                                  uint32_t* nlist,
                                  uint32_t* next_mask)
{
    MI_GRID_STRIDE(tid, N)
    {
        const int32_t idx = p2r[tid];
        if(idx < 0)
            continue;
        const int32_t cidx = cprev[idx];
        if(cidx < 0 || cidx >= N)
            continue;
        uint32_t       conn = 0, nk = 0;
        const uint32_t st = nstart[idx], cnt = nnum[idx];
        for(uint32_t q = 0; q < cnt; ++q)
        {
            const uint32_t w  = nlist[st + q];
            const int32_t  cc = cprev[w];
            if(cc < 0 || cc >= N)
            {
                nlist[st + nk++] = w;
                continue;
            }
            if(cidx / kMasBank == cc / kMasBank)
                conn |= 1u << (cc % kMasBank);
            else
                nlist[st + nk++] = w;
        }
        nnum[idx] = nk;
        if(conn)
            atomicOr(&next_mask[cidx], conn);
    }
}

// next_level_cluster (296-338)
__global__ void lx_cluster_kernel(int32_t nbx, int32_t Nx, uint32_t* next_mask, uint32_t* pref)
{
    MI_GRID_STRIDE(b, nbx)
    {
        uint32_t cache[kMasBank];
        for(int l = 0; l < kMasBank; ++l)
        {
            const int64_t i = b * kMasBank + l;
            cache[l]        = i < Nx ? ((1u << l) | next_mask[i]) : 0u;
        }
        uint32_t cnt = 0;
        for(int l = 0; l < kMasBank; ++l)
        {
            const int64_t i = b * kMasBank + l;
            if(i >= Nx)
                continue;
            const uint32_t c = closure16(cache, l, cache[l]);
            next_mask[i]     = c;
            cnt += __popc(c & lt16(l)) == 0;
        }
        pref[b] = cnt;
    }
}

// prefix_sum_Lx (340-388)
__global__ void lx_prefix_kernel(int32_t nbx,
                                 int32_t Nx,
                                 int32_t begin,
                                 int32_t next_begin,
                                 uint32_t* next_mask,
                                 const uint32_t* __restrict__ psum,
                                 int32_t* going_next)
{
    MI_GRID_STRIDE(b, nbx)
    {
        uint32_t elected = 0;
        for(int l = 0; l < kMasBank; ++l)
        {
            const int64_t i = b * kMasBank + l;
            if(i < Nx && __popc(next_mask[i] & lt16(l)) == 0)
                elected |= 1u << l;
        }
        int32_t nv[kMasBank];
        for(int l = 0; l < kMasBank; ++l)
        {
            const int64_t i = b * kMasBank + l;
            if(i >= Nx)
                continue;
            const int el          = __ffs(next_mask[i]) - 1;
            nv[l]                 = __popc(elected & lt16(el)) + (int32_t)psum[b];
            going_next[begin + i] = nv[l] + next_begin;
        }
        for(int l = 0; l < kMasBank; ++l)
        {
            const int64_t i = b * kMasBank + l;
            if(i < Nx)
                next_mask[i] = (uint32_t)nv[l];
        }
    }
}

// compute_next_level (390-409)
__global__ void lx_next_kernel(int32_t V, const int32_t* __restrict__ cprev, const uint32_t* __restrict__ next_mask,
                               int32_t* ccur)
{
    MI_GRID_STRIDE(idx, V)
    {
        const int32_t nx = cprev[idx];
        ccur[idx]        = (nx < 0 || nx >= V) ? -1 : (int32_t)next_mask[nx];
    }
}

// aggregation_kernel (411-466): per vertex, its node at every level
__global__ void aggregate_kernel(int32_t V, int L, const int32_t* __restrict__ going_next, int64_t gn_size,
                                 int32_t* table)
{
    MI_GRID_STRIDE(idx, V)
    {
        const int32_t first = going_next[idx];
        if(first < 0)
            continue;
        int32_t* tab = table + (int64_t)kMasMaxLvl * idx;
        tab[0]       = first;
        int32_t cur  = first;
        for(int l = 1; l < L - 1; ++l)
        {
            if(cur < 0 || cur >= gn_size)
                break;
            cur    = going_next[cur];
            tab[l] = cur;
        }
    }
}

// Every copy below is ordered on the stream the kernels run on and waited
// for: a plain hipMemcpy runs on the legacy null stream, which does not wait
// for work on a non-blocking stream.

// a device array of the descriptor, or the host array uploaded into `tmp`
const int32_t* device_input(const int32_t* dev, const int32_t* host, int64_t n, DBuf<int32_t>& tmp, hipStream_t s)
{
    if(dev || n <= 0)
        return dev;
    MI_REQUIRE(host != nullptr, MI_ERR_ARG, "MAS: a topology array is NULL with a nonzero count");
    tmp.resize((size_t)n);
    MI_HIP(hipMemcpyAsync(tmp.data(), host, sizeof(int32_t) * (size_t)n, hipMemcpyHostToDevice, s));
    MI_HIP(hipStreamSynchronize(s));
    return tmp.data();
}

template <typename T>
T read_back(const T* d, hipStream_t s)
{
    T v;
    MI_HIP(hipMemcpyAsync(&v, d, sizeof(T), hipMemcpyDeviceToHost, s));
    MI_HIP(hipStreamSynchronize(s));
    return v;
}

template <typename T>
void copy_back(std::vector<T>& h, const DBuf<T>& d, size_t n, hipStream_t s)
{
    h.resize(n);
    if(n)
    {
        MI_HIP(hipMemcpyAsync(h.data(), d.data(), sizeof(T) * n, hipMemcpyDeviceToHost, s));
        MI_HIP(hipStreamSynchronize(s));
    }
}
}  // namespace

void build_adjacency_device(int32_t          V,
                            const int32_t*   tets,
                            int64_t          n_tets,
                            const int32_t*   tris,
                            int64_t          n_tris,
                            const int32_t*   edges,
                            int64_t          n_edges,
                            DeviceAdjacency& out,
                            hipStream_t      s)
{
    const int64_t nt = tets ? std::max<int64_t>(0, n_tets) : 0, ntr = tris ? std::max<int64_t>(0, n_tris) : 0,
                  ne = edges ? std::max<int64_t>(0, n_edges) : 0;
    const int64_t E  = 12 * nt + 6 * ntr + 2 * ne;
    MI_REQUIRE(E < ((int64_t)1 << 31), MI_ERR_ARG, "adjacency: more than 2^31-1 vertex pairs");
    out.nstart.resize((size_t)std::max(V, 1));
    out.nnum.resize((size_t)std::max(V, 1));
    MI_HIP(hipMemsetAsync(out.nnum.data(), 0, out.nnum.bytes(), s));
    out.M = 0;
    DBuf<uint32_t> chunk;
    if(E > 0 && V > 0)
    {
        DBuf<uint64_t> keys[2];
        DBuf<int32_t>  vals[2];
        DBuf<uint32_t> hist, head;
        DBuf<int64_t>  cnt;
        keys[0].resize((size_t)E);
        keys[1].resize((size_t)E);
        vals[0].resize((size_t)E);
        vals[1].resize((size_t)E);
        head.resize((size_t)E);
        cnt.resize(2);
        adj_emit_kernel<<<grid_n(nt + ntr + ne), kT, 0, s>>>(tets, nt, tris, ntr, edges, ne, V, keys[0].data());
        MI_HIP(hipMemcpyAsync(cnt.data(), &E, sizeof(int64_t), hipMemcpyHostToDevice, s));
        const int bits = radix_key_bits((uint64_t)V * (uint64_t)V);
        const int cur  = radix_sort_pairs(keys[0].data(), vals[0].data(), keys[1].data(), vals[1].data(), cnt.data(),
                                          E, bits, hist, chunk, s);
        adj_head_kernel<<<grid_n(E), kT, 0, s>>>(keys[cur].data(), E, V, head.data(), out.nnum.data());
        exclusive_scan_u32(head.data(), E, chunk, cnt.data() + 1, s);
        out.M = read_back(cnt.data() + 1, s);
        out.nlist.resize((size_t)std::max<int64_t>(out.M, 1));
        adj_write_kernel<<<grid_n(E), kT, 0, s>>>(keys[cur].data(), E, V, head.data(), out.nlist.data());
    }
    else
        out.nlist.resize(1);
    if(V > 0)
    {
        MI_HIP(hipMemcpyAsync(out.nstart.data(), out.nnum.data(), sizeof(uint32_t) * V, hipMemcpyDeviceToDevice, s));
        exclusive_scan_u32(out.nstart.data(), V, chunk, nullptr, s);
    }
    MI_HIP(hipGetLastError());
    MI_HIP(hipStreamSynchronize(s));
}

MasHierarchy mas_build_hierarchy_device(const mi_precond_desc& d)
{
    MasHierarchy  h;
    const int32_t V = d.vert_num;
    h.V             = V;
    MI_REQUIRE(V >= 0 && (int64_t)V * V < ((int64_t)1 << 62), MI_ERR_ARG, "MAS: vertex count out of range");
    // order after any work the caller queued for the device inputs
    MI_HIP(hipDeviceSynchronize());
    const hipStream_t s = nullptr;
    DBuf<int32_t>     t_tets, t_tris, t_edges, t_part;
    const int32_t*    tets  = device_input(d.tets_dev, d.tets_host, 4 * d.n_tets, t_tets, s);
    const int32_t*    tris  = device_input(d.tris_dev, d.tris_host, 3 * d.n_tris, t_tris, s);
    const int32_t*    edges = device_input(d.edges_dev, d.edges_host, 2 * d.n_edges, t_edges, s);
    const int32_t*    part  = device_input(d.part_ids_dev, d.part_ids_host, V, t_part, s);
    MI_REQUIRE(V == 0 || part, MI_ERR_ARG, "MAS: partition ids are required");

    // ---- adjacency CSR
    DeviceAdjacency adj;
    build_adjacency_device(V, tets, d.n_tets, tris, d.n_tris, edges, d.n_edges, adj, s);
    DBuf<uint32_t>& nstart = adj.nstart;
    DBuf<uint32_t>& nnum   = adj.nnum;
    DBuf<uint32_t>& nlist  = adj.nlist;

    // ---- partition slots
    DBuf<int32_t> stat;
    stat.resize(3);
    MI_HIP(hipMemsetAsync(stat.data(), 0, stat.bytes(), s));
    if(V > 0)
        part_stat_kernel<<<grid_n(V), kT, 0, s>>>(part, V, stat.data());
    int32_t st[3];
    MI_HIP(hipMemcpyAsync(st, stat.data(), sizeof(st), hipMemcpyDeviceToHost, s));
    MI_HIP(hipStreamSynchronize(s));
    const int32_t  max_part = st[0], P = max_part + 1;
    h.has_unpart            = st[1] != 0;
    DBuf<uint32_t> psize, poff, chunk;
    DBuf<int32_t>  members;
    DBuf<int64_t>  tot;
    psize.resize((size_t)P);
    poff.resize((size_t)P);
    members.resize((size_t)kMasBank * P);
    tot.resize(1);
    MI_HIP(hipMemsetAsync(psize.data(), 0, psize.bytes(), s));
    if(V > 0)
        part_collect_kernel<<<grid_n(V), kT, 0, s>>>(part, V, psize.data(), members.data(), stat.data());
    MI_HIP(hipMemcpyAsync(st, stat.data(), sizeof(st), hipMemcpyDeviceToHost, s));
    MI_HIP(hipStreamSynchronize(s));
    MI_REQUIRE(st[2] == 0, MI_ERR_ARG, "MAS: a partition holds more than 16 vertices");
    part_align_kernel<<<grid_n(P), kT, 0, s>>>(psize.data(), P, poff.data());
    exclusive_scan_u32(poff.data(), P, chunk, tot.data(), s);
    const int64_t N64 = read_back(tot.data(), s);
    MI_REQUIRE(N64 < ((int64_t)1 << 31), MI_ERR_ARG, "MAS: too many partition slots");
    const int32_t N = (int32_t)N64;
    h.N             = N;
    DBuf<int32_t> p2r, r2p;
    p2r.resize((size_t)std::max(N, 1));
    r2p.resize((size_t)std::max(V, 1));
    fill_value_kernel<int32_t><<<grid_n(p2r.size()), kT, 0, s>>>(p2r.data(), (int64_t)p2r.size(), -1);
    fill_value_kernel<int32_t><<<grid_n(r2p.size()), kT, 0, s>>>(r2p.data(), (int64_t)r2p.size(), -1);
    part_fill_kernel<<<grid_n(P), kT, 0, s>>>(psize.data(), poff.data(), members.data(), P, p2r.data(), r2p.data());

    // ---- levels (init_neighbor, mas_preconditioner_engine.cu:1974-2024)
    const int max_nodes = std::max(N, V);
    h.padded            = mas_bank_align(max_nodes);
    h.L                 = mas_num_levels(max_nodes);
    const int L         = h.L;
    h.level_count.assign((size_t)L + 1, 0);
    h.level_begin.assign((size_t)L + 1, 0);
    const int64_t  gn_size = (int64_t)h.padded * L;
    DBuf<int32_t>  gn, coarse, table;
    DBuf<uint32_t> fine, pref, mask;
    gn.resize((size_t)std::max<int64_t>(gn_size, 1));
    coarse.resize((size_t)std::max<int64_t>((int64_t)V * L, 1));
    table.resize((size_t)std::max<int64_t>((int64_t)V * kMasMaxLvl, 1));
    fine.resize((size_t)std::max(max_nodes, 1));
    mask.resize((size_t)std::max(max_nodes, 1));
    fill_value_kernel<int32_t><<<grid_n(gn.size()), kT, 0, s>>>(gn.data(), (int64_t)gn.size(), -1);
    fill_value_kernel<int32_t><<<grid_n(coarse.size()), kT, 0, s>>>(coarse.data(), (int64_t)coarse.size(), -1);
    fill_value_kernel<int32_t><<<grid_n(table.size()), kT, 0, s>>>(table.data(), (int64_t)table.size(), -1);
    fill_value_kernel<uint32_t><<<grid_n(fine.size()), kT, 0, s>>>(fine.data(), (int64_t)fine.size(), 0u);
    const int nbanks = N / kMasBank;
    pref.resize((size_t)std::max(max_nodes / kMasBank + 2, nbanks + 1));
    if(N >= 1)
    {
        l0_connect_kernel<<<grid_n(N), kT, 0, s>>>(N, p2r.data(), r2p.data(), nstart.data(), nnum.data(),
                                                   nlist.data(), fine.data());
        l0_prefix_kernel<<<grid_n(nbanks), kT, 0, s>>>(nbanks, p2r.data(), fine.data(), pref.data());
        exclusive_scan_u32(pref.data(), nbanks, chunk, tot.data(), s);
        const int level1_begin = mas_bank_align(std::max(mas_bank_align(N), V));
        h.level1_begin         = level1_begin;
        level1_kernel<<<grid_n(nbanks), kT, 0, s>>>(nbanks, p2r.data(), fine.data(), pref.data(), level1_begin,
                                                    coarse.data(), gn.data());
        h.level_count[1] = (int32_t)read_back(tot.data(), s);
        h.level_begin[1] = level1_begin;
        for(int level = 1; level < L; ++level)
        {
            const int32_t* cprev = coarse.data() + (int64_t)(level - 1) * V;
            int32_t*       ccur  = coarse.data() + (int64_t)level * V;
            fill_value_kernel<uint32_t><<<grid_n(mask.size()), kT, 0, s>>>(mask.data(), (int64_t)mask.size(), 0u);
            lx_connect_kernel<<<grid_n(N), kT, 0, s>>>(N, p2r.data(), cprev, nstart.data(), nnum.data(),
                                                       nlist.data(), mask.data());
            const int32_t Nx = h.level_count[level], begin = h.level_begin[level];
            const int32_t nbx = (Nx + kMasBank - 1) / kMasBank;
            if(nbx > 0)
            {
                lx_cluster_kernel<<<grid_n(nbx), kT, 0, s>>>(nbx, Nx, mask.data(), pref.data());
                exclusive_scan_u32(pref.data(), nbx, chunk, tot.data(), s);
            }
            else
                MI_HIP(hipMemsetAsync(tot.data(), 0, sizeof(int64_t), s));
            const int32_t next_begin = begin + mas_bank_align(Nx);
            if(nbx > 0)
                lx_prefix_kernel<<<grid_n(nbx), kT, 0, s>>>(nbx, Nx, begin, next_begin, mask.data(), pref.data(),
                                                            gn.data());
            h.level_count[level + 1] = (int32_t)read_back(tot.data(), s);
            h.level_begin[level + 1] = next_begin;
            if(V > 0)
                lx_next_kernel<<<grid_n(V), kT, 0, s>>>(V, cprev, mask.data(), ccur);
        }
        h.total = h.level_begin[L];
        if(V > 0)
            aggregate_kernel<<<grid_n(V), kT, 0, s>>>(V, L, gn.data(), gn_size, table.data());
    }
    MI_HIP(hipGetLastError());
    MI_HIP(hipStreamSynchronize(s));
    copy_back(h.part_to_real, p2r, (size_t)std::max(N, 1), s);
    copy_back(h.real_to_part, r2p, (size_t)V, s);
    copy_back(h.going_next, gn, (size_t)gn_size, s);
    copy_back(h.table, table, (size_t)V * kMasMaxLvl, s);
    return h;
}
}  // namespace mi
