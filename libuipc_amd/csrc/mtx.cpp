// Matrix Market readers (host C++), the counterpart of the reference's debug
// dumps (utils/matrix_market.h:26-158).  The writers live in linsys.hip next to
// the assembled matrix they serialise.
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../../include/mi_linsys.h"

namespace mi
{
void set_last_error(const std::string& msg);
}

namespace
{
struct Block
{
    double v[9];
};

bool read_header(FILE* f, std::string& banner, char* line, size_t cap)
{
    if(!std::fgets(line, (int)cap, f))
        return false;
    banner = line;
    do
    {
        if(!std::fgets(line, (int)cap, f))
            return false;
    } while(line[0] == '%');
    return true;
}
}  // namespace

extern "C" int mi_mtx_read_block_triplets(const char* path,
                                          int         one_based,
                                          int32_t*    block_rows_out,
                                          int64_t*    n_triplets_out,
                                          int32_t*    rows_host,
                                          int32_t*    cols_host,
                                          double*     vals_host)
{
    if(!path || !n_triplets_out)
    {
        mi::set_last_error("mi_mtx_read_block_triplets: null argument");
        return MI_ERR_ARG;
    }
    FILE* f = std::fopen(path, "r");
    if(!f)
    {
        mi::set_last_error(std::string("mi_mtx_read_block_triplets: cannot open ") + path);
        return MI_ERR_ARG;
    }
    char        line[512];
    std::string banner;
    if(!read_header(f, banner, line, sizeof(line)) || banner.find("coordinate") == std::string::npos)
    {
        std::fclose(f);
        mi::set_last_error("mi_mtx_read_block_triplets: not a coordinate Matrix Market file");
        return MI_ERR_ARG;
    }
    long long nr = 0, nc = 0, nnz = 0;
    if(std::sscanf(line, "%lld %lld %lld", &nr, &nc, &nnz) != 3 || nr != nc || nr % 3 != 0 || nr < 0)
    {
        std::fclose(f);
        mi::set_last_error("mi_mtx_read_block_triplets: size line must be 'n n nnz' with n % 3 == 0");
        return MI_ERR_ARG;
    }
    const int64_t           B   = nr / 3;
    const long long         off = one_based ? 1 : 0;
    std::map<uint64_t, Block> blocks;
    for(long long e = 0; e < nnz; ++e)
    {
        long long i, j;
        double    v;
        if(std::fscanf(f, "%lld %lld %lf", &i, &j, &v) != 3)
        {
            std::fclose(f);
            mi::set_last_error("mi_mtx_read_block_triplets: truncated entry list");
            return MI_ERR_ARG;
        }
        i -= off;
        j -= off;
        if(i < 0 || j < 0 || i >= nr || j >= nc)
        {
            std::fclose(f);
            mi::set_last_error("mi_mtx_read_block_triplets: entry outside the matrix");
            return MI_ERR_ARG;
        }
        uint64_t key = (uint64_t)(i / 3) * (uint64_t)B + (uint64_t)(j / 3);
        auto     it  = blocks.find(key);
        if(it == blocks.end())
        {
            Block z;
            std::memset(z.v, 0, sizeof(z.v));
            it = blocks.emplace(key, z).first;
        }
        it->second.v[(i % 3) + 3 * (j % 3)] += v;
    }
    std::fclose(f);
    if(block_rows_out)
        *block_rows_out = (int32_t)B;
    *n_triplets_out = (int64_t)blocks.size();
    if(rows_host && cols_host && vals_host)
    {
        int64_t t = 0;
        for(auto& kv : blocks)
        {
            rows_host[t] = (int32_t)(kv.first / (uint64_t)B);
            cols_host[t] = (int32_t)(kv.first % (uint64_t)B);
            std::memcpy(vals_host + 9 * t, kv.second.v, sizeof(kv.second.v));
            ++t;
        }
    }
    mi::set_last_error("");
    return MI_OK;
}

extern "C" int mi_vec_read_mtx(const char* path, int64_t* n_out, double* x_host)
{
    if(!path || !n_out)
    {
        mi::set_last_error("mi_vec_read_mtx: null argument");
        return MI_ERR_ARG;
    }
    FILE* f = std::fopen(path, "r");
    if(!f)
    {
        mi::set_last_error(std::string("mi_vec_read_mtx: cannot open ") + path);
        return MI_ERR_ARG;
    }
    char        line[512];
    std::string banner;
    long long   n = 0, ncol = 0;
    if(!read_header(f, banner, line, sizeof(line)) || banner.find("array") == std::string::npos
       || std::sscanf(line, "%lld %lld", &n, &ncol) != 2 || ncol != 1 || n < 0)
    {
        std::fclose(f);
        mi::set_last_error("mi_vec_read_mtx: expected an 'array' Matrix Market file with one column");
        return MI_ERR_ARG;
    }
    *n_out = n;
    if(x_host)
        for(long long i = 0; i < n; ++i)
            if(std::fscanf(f, "%lf", &x_host[i]) != 1)
            {
                std::fclose(f);
                mi::set_last_error("mi_vec_read_mtx: truncated value list");
                return MI_ERR_ARG;
            }
    std::fclose(f);
    mi::set_last_error("");
    return MI_OK;
}
