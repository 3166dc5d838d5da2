// Seeded synthetic linear systems in libuipc's triplet format (host C++).
//
// What is generated (one Newton system of an implicit-Euler FEM step):
//   A = M + dt^2 (V * d2Psi/dx2 projected to PSD) + dt^2 * sum_contacts kappa g g^T
// which is the reference's scaling (kinetic block m_i I,
// finite_element/bdf/finite_element_bdf1_kinetic.cu:60-70; elastic energy
// scaled by V dt^2, finite_element/constitutions/stable_neo_hookean_3d.cu:51,106).
//
// Triplet order mirrors FEMLinearSubsystem's layout
// (finite_element/fem_linear_subsystem.cu:276-355): [kinetic, one diagonal block
// per vertex | 10 half-blocks per tet | 10 half-blocks per contact stencil].
// A half-block write keeps the GLOBAL row < col by swapping the stencil slots
// and writing that slot pair's block, exactly as
// TripletMatrixAssembler::ProxyRangeHalf::write does (utils/matrix_assembler.h:312-410).
//
// Energy: stable Neo-Hookean  Psi = mu/2 (tr(F^T F) - 3) + lambda/2 (J - 1 - mu/lambda)^2
// (the form the reference's SNK1 path evaluates, detail/stable_neo_hookean_3d_snk1.inl),
// Hessian projected to PSD by a Jacobi eigen-decomposition of the 9x9 dP/dF.
//
// Randomness is counter based (splitmix64 of seed and an index), so the output
// does not depend on the thread count.
#include "../../include/mi_synth.h"

#include <algorithm>
#include <functional>
#include <atomic>
#include <cmath>
#include <cstring>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

namespace
{
thread_local std::string g_err;

inline uint64_t splitmix64(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

// deterministic stream: value k of stream (seed, tag, index)
struct Rng
{
    uint64_t state;
    Rng(uint64_t seed, uint64_t tag, uint64_t index)
        : state(splitmix64(seed ^ splitmix64(tag * 0x632BE59BD9B4E019ull + index)))
    {
    }
    uint64_t next()
    {
        state += 0x9E3779B97F4A7C15ull;
        uint64_t z = state;
        z          = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z          = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform01() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
    double uniform(double lo, double hi) { return lo + (hi - lo) * uniform01(); }
    double normal()
    {
        double u1 = uniform01();
        double u2 = uniform01();
        if(u1 < 1e-300)
            u1 = 1e-300;
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
    int64_t below(int64_t n) { return (int64_t)(next() % (uint64_t)n); }
};

enum : uint64_t
{
    TAG_JITTER  = 1,
    TAG_RHS     = 2,
    TAG_CONTACT = 3,
};

struct Grid
{
    int     N;
    int64_t V;
    int64_t cells;
    int64_t tets;
    int64_t vid(int i, int j, int k) const { return i + (int64_t)N * (j + (int64_t)N * k); }
};

// The 5-tet split of a cube cell, alternating with the parity of i+j+k so the
// face diagonals of neighbouring cells agree.  Corner code = a | b<<1 | c<<2.
const int kTetsEven[5][4] = {{1, 0, 3, 5}, {2, 0, 6, 3}, {4, 0, 5, 6}, {7, 3, 6, 5}, {0, 3, 5, 6}};
const int kTetsOdd[5][4]  = {{0, 1, 2, 4}, {3, 1, 7, 2}, {5, 1, 4, 7}, {6, 2, 7, 4}, {1, 2, 4, 7}};

void cell_tet(const Grid& g, int64_t t, int32_t out[4])
{
    int64_t c  = t / 5;
    int     s  = (int)(t % 5);
    int     nc = g.N - 1;
    int     i  = (int)(c % nc);
    int     j  = (int)((c / nc) % nc);
    int     k  = (int)(c / ((int64_t)nc * nc));
    const int(*tab)[4] = ((i + j + k) & 1) ? kTetsOdd : kTetsEven;
    for(int q = 0; q < 4; ++q)
    {
        int code = tab[s][q];
        out[q]   = (int32_t)g.vid(i + (code & 1), j + ((code >> 1) & 1), k + ((code >> 2) & 1));
    }
}

inline double det3(const double m[9])  // column-major
{
    return m[0] * (m[4] * m[8] - m[7] * m[5]) - m[3] * (m[1] * m[8] - m[7] * m[2])
           + m[6] * (m[1] * m[5] - m[4] * m[2]);
}

void inv3(const double m[9], double out[9])
{
    double d   = det3(m);
    double id  = 1.0 / d;
    out[0]     = (m[4] * m[8] - m[7] * m[5]) * id;
    out[3]     = -(m[3] * m[8] - m[6] * m[5]) * id;
    out[6]     = (m[3] * m[7] - m[6] * m[4]) * id;
    out[1]     = -(m[1] * m[8] - m[7] * m[2]) * id;
    out[4]     = (m[0] * m[8] - m[6] * m[2]) * id;
    out[7]     = -(m[0] * m[7] - m[6] * m[1]) * id;
    out[2]     = (m[1] * m[5] - m[4] * m[2]) * id;
    out[5]     = -(m[0] * m[5] - m[3] * m[2]) * id;
    out[8]     = (m[0] * m[4] - m[3] * m[1]) * id;
}

// symmetric Jacobi eigen-decomposition, A (n x n row-major, destroyed),
// eigenvalues w, eigenvectors as columns of Q (row-major)
template <int n>
void jacobi_eigen(double A[n][n], double w[n], double Q[n][n])
{
    for(int i = 0; i < n; ++i)
        for(int j = 0; j < n; ++j)
            Q[i][j] = (i == j) ? 1.0 : 0.0;
    double scale = 0.0;
    for(int i = 0; i < n; ++i)
        for(int j = 0; j < n; ++j)
            scale = std::max(scale, std::fabs(A[i][j]));
    if(scale == 0.0)
    {
        for(int i = 0; i < n; ++i)
            w[i] = 0.0;
        return;
    }
    for(int sweep = 0; sweep < 50; ++sweep)
    {
        double off = 0.0;
        for(int p = 0; p < n; ++p)
            for(int q = p + 1; q < n; ++q)
                off += A[p][q] * A[p][q];
        if(off <= 1e-30 * scale * scale)
            break;
        for(int p = 0; p < n; ++p)
            for(int q = p + 1; q < n; ++q)
            {
                double apq = A[p][q];
                if(std::fabs(apq) <= 1e-300)
                    continue;
                double theta = (A[q][q] - A[p][p]) / (2.0 * apq);
                double t     = (theta >= 0 ? 1.0 : -1.0)
                           / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
                double c = 1.0 / std::sqrt(t * t + 1.0);
                double s = t * c;
                for(int k = 0; k < n; ++k)
                {
                    double akp = A[k][p], akq = A[k][q];
                    A[k][p]    = c * akp - s * akq;
                    A[k][q]    = s * akp + c * akq;
                }
                for(int k = 0; k < n; ++k)
                {
                    double apk = A[p][k], aqk = A[q][k];
                    A[p][k]    = c * apk - s * aqk;
                    A[q][k]    = s * apk + c * aqk;
                }
                for(int k = 0; k < n; ++k)
                {
                    double qkp = Q[k][p], qkq = Q[k][q];
                    Q[k][p]    = c * qkp - s * qkq;
                    Q[k][q]    = s * qkp + c * qkq;
                }
            }
    }
    for(int i = 0; i < n; ++i)
        w[i] = A[i][i];
}

inline void cross3(const double a[3], const double b[3], double o[3])
{
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}

// skew(v) * u = v x u ; writes +s*skew(v) into the 3x3 block (row-major dst with stride 9)
inline void add_skew(double H[9][9], int br, int bc, const double v[3], double s)
{
    // skew(v) = [[0,-v2,v1],[v2,0,-v0],[-v1,v0,0]]
    H[3 * br + 0][3 * bc + 1] += -s * v[2];
    H[3 * br + 0][3 * bc + 2] += s * v[1];
    H[3 * br + 1][3 * bc + 0] += s * v[2];
    H[3 * br + 1][3 * bc + 2] += -s * v[0];
    H[3 * br + 2][3 * bc + 0] += -s * v[1];
    H[3 * br + 2][3 * bc + 1] += s * v[0];
}

// 9x9 PSD-projected d2Psi/dF2 in vec(F) column-major order (index a + 3 b)
void snh_hessian_psd(const double F[9], double mu, double lambda, double Hp[9][9])
{
    const double* f0 = F + 0;
    const double* f1 = F + 3;
    const double* f2 = F + 6;
    double        g[9];
    cross3(f1, f2, g + 0);
    cross3(f2, f0, g + 3);
    cross3(f0, f1, g + 6);
    double J     = det3(F);
    double alpha = 1.0 + mu / lambda;
    double H[9][9];
    for(int i = 0; i < 9; ++i)
        for(int j = 0; j < 9; ++j)
            H[i][j] = (i == j ? mu : 0.0) + lambda * g[i] * g[j];
    double s = lambda * (J - alpha);
    // d2J/dF2: block(0,1) = -skew(f2), block(0,2) = skew(f1), block(1,2) = -skew(f0), symmetric
    add_skew(H, 0, 1, f2, -s);
    add_skew(H, 1, 0, f2, s);
    add_skew(H, 0, 2, f1, s);
    add_skew(H, 2, 0, f1, -s);
    add_skew(H, 1, 2, f0, -s);
    add_skew(H, 2, 1, f0, s);

    double w[9], Q[9][9];
    jacobi_eigen<9>(H, w, Q);
    for(int i = 0; i < 9; ++i)
        for(int j = 0; j < 9; ++j)
        {
            double acc = 0.0;
            for(int k = 0; k < 9; ++k)
                if(w[k] > 0.0)
                    acc += Q[i][k] * w[k] * Q[j][k];
            Hp[i][j] = acc;
        }
    // exact symmetry
    for(int i = 0; i < 9; ++i)
        for(int j = i + 1; j < 9; ++j)
        {
            double a = 0.5 * (Hp[i][j] + Hp[j][i]);
            Hp[i][j] = Hp[j][i] = a;
        }
}

// writes the 10 half-block triplets of a 4-node stencil whose 12x12 local
// matrix is K (row-major, rows (node, comp))
inline void emit_half4(const int32_t idx[4], const double K[12][12], int32_t* rows, int32_t* cols, double* vals)
{
    int o = 0;
    for(int ii = 0; ii < 4; ++ii)
        for(int jj = ii; jj < 4; ++jj)
        {
            int L = ii, R = jj;
            if(!(idx[ii] < idx[jj]))
            {
                L = jj;
                R = ii;
            }
            rows[o] = idx[L];
            cols[o] = idx[R];
            double* v = vals + 9 * o;
            for(int c = 0; c < 3; ++c)
                for(int r = 0; r < 3; ++r)
                    v[r + 3 * c] = K[3 * L + r][3 * R + c];
            ++o;
        }
}

// contact stencil s: a random vertex p and a face (3 vertices, p excluded when
// possible) of a tet in a cell within contact_cells of p; g ~ N(0,1)^12 and
// kappa log-uniform in [kappa_lo, kappa_hi]
void contact_stencil(const mi_synth_fem_desc* d, const Grid& g, const int32_t* tv, int64_t s, int32_t idx[4],
                     double gv[12], double& kappa)
{
    const int32_t nc = g.N - 1;
    Rng     rng(d->seed, TAG_CONTACT, (uint64_t)s);
    int64_t p  = rng.below(g.V);
    int     pi = (int)(p % g.N), pj = (int)((p / g.N) % g.N), pk = (int)(p / ((int64_t)g.N * g.N));
    int     R  = std::max(1, d->contact_cells);
    auto    pick = [&](int pc)
    {
        int lo = std::max(0, pc - R), hi = std::min(nc - 1, pc + R - 1);
        if(hi < lo)
            hi = lo;
        return lo + (int)rng.below(hi - lo + 1);
    };
    int            ci = pick(pi), cj = pick(pj), ck = pick(pk);
    int64_t        cell = ci + (int64_t)nc * (cj + (int64_t)nc * ck);
    int            sub  = (int)rng.below(5);
    int            f0   = (int)rng.below(4);
    const int32_t* tq   = &tv[4 * (cell * 5 + sub)];
    idx[0]              = (int32_t)p;
    for(int f = 0; f < 4; ++f)
    {
        int  omit  = (f0 + f) % 4;  // face = the 3 tet vertices except `omit`
        int  o     = 1;
        bool has_p = false;
        for(int c = 0; c < 4; ++c)
            if(c != omit)
            {
                idx[o++] = tq[c];
                has_p |= (tq[c] == (int32_t)p);
            }
        if(!has_p)
            break;
    }
    for(int c = 0; c < 12; ++c)
        gv[c] = rng.normal();
    kappa = d->kappa_lo * std::pow(d->kappa_hi / d->kappa_lo, rng.uniform01());
}

void parallel_for(int64_t n, int threads, const std::function<void(int64_t, int64_t)>& fn);

int hw_threads(int requested)
{
    if(requested > 0)
        return requested;
    unsigned h = std::thread::hardware_concurrency();
    return h ? (int)std::min(h, 64u) : 4;
}

void parallel_for(int64_t n, int threads, const std::function<void(int64_t, int64_t)>& fn)
{
    if(n <= 0)
        return;
    int T = (int)std::min<int64_t>(threads, std::max<int64_t>(1, n / 1024));
    if(T <= 1)
    {
        fn(0, n);
        return;
    }
    std::vector<std::thread> pool;
    std::atomic<int64_t>     next{0};
    const int64_t            chunk = std::max<int64_t>(1024, n / (T * 16));
    for(int t = 0; t < T; ++t)
        pool.emplace_back(
            [&]()
            {
                while(true)
                {
                    int64_t b = next.fetch_add(chunk);
                    if(b >= n)
                        break;
                    fn(b, std::min(n, b + chunk));
                }
            });
    for(auto& th : pool)
        th.join();
}

uint32_t part1by2(uint32_t x)  // spread 10 bits to every third bit
{
    x &= 0x3ff;
    x = (x | (x << 16)) & 0x030000FF;
    x = (x | (x << 8)) & 0x0300F00F;
    x = (x | (x << 4)) & 0x030C30C3;
    x = (x | (x << 2)) & 0x09249249;
    return x;
}

// rest positions X (the regular grid), deformed positions x (jittered), tets
// (oriented by rest volume) and the lumped vertex masses
void build_mesh(const mi_synth_fem_desc* d,
                const Grid&              g,
                int                      T,
                std::vector<double>&     X,
                std::vector<double>&     x,
                std::vector<int32_t>&    tv,
                std::vector<double>&     mass)
{
    const double h = 1.0 / (double)(g.N - 1);
        // rest and deformed positions
        X.assign(3 * g.V, 0.0);
        x.assign(3 * g.V, 0.0);
        parallel_for(g.V,
                     T,
                     [&](int64_t b, int64_t e)
                     {
                         for(int64_t v = b; v < e; ++v)
                         {
                             int i = (int)(v % g.N), j = (int)((v / g.N) % g.N),
                                 k = (int)(v / ((int64_t)g.N * g.N));
                             Rng rng(d->seed, TAG_JITTER, (uint64_t)v);
                             X[3 * v + 0] = i * h;
                             X[3 * v + 1] = j * h;
                             X[3 * v + 2] = k * h;
                             for(int c = 0; c < 3; ++c)
                                 x[3 * v + c] = X[3 * v + c] + d->jitter * h * rng.uniform(-1.0, 1.0);
                         }
                     });

        // tets (oriented by rest volume)
        tv.assign(4 * g.tets, 0);
        parallel_for(g.tets,
                     T,
                     [&](int64_t b, int64_t e)
                     {
                         for(int64_t t = b; t < e; ++t)
                         {
                             int32_t q[4];
                             cell_tet(g, t, q);
                             double Dm[9];
                             for(int c = 0; c < 3; ++c)
                                 for(int r = 0; r < 3; ++r)
                                     Dm[r + 3 * c] = X[3 * q[c + 1] + r] - X[3 * q[0] + r];
                             if(det3(Dm) < 0)
                                 std::swap(q[2], q[3]);
                             for(int c = 0; c < 4; ++c)
                                 tv[4 * t + c] = q[c];
                         }
                     });

        // lumped mass: every tet of the regular grid has rest volume h^3 / 6
        // or h^3 / 3 (the central tet); accumulate per vertex exactly
        mass.assign(g.V, 0.0);
        for(int64_t t = 0; t < g.tets; ++t)
        {
            const int32_t* q = &tv[4 * t];
            double         Dm[9];
            for(int c = 0; c < 3; ++c)
                for(int r = 0; r < 3; ++r)
                    Dm[r + 3 * c] = X[3 * q[c + 1] + r] - X[3 * q[0] + r];
            double m = d->density * std::fabs(det3(Dm)) / 6.0 / 4.0;
            for(int c = 0; c < 4; ++c)
                mass[q[c]] += m;
        }

}

int check_desc(const mi_synth_fem_desc* d)
{
    if(!d)
        return (g_err = "mi_synth: desc is NULL", -1);
    if(d->grid < 2 || d->grid > 1024)
        return (g_err = "mi_synth: grid must be in [2, 1024]", -1);
    if(d->n_contacts < 0)
        return (g_err = "mi_synth: n_contacts < 0", -1);
    if(d->n_contacts > 0 && (d->kappa_lo <= 0 || d->kappa_hi < d->kappa_lo))
        return (g_err = "mi_synth: invalid kappa range", -1);
    if(d->youngs <= 0 || d->poisson <= -1.0 || d->poisson >= 0.5 || d->density <= 0 || d->dt <= 0)
        return (g_err = "mi_synth: invalid material or dt", -1);
    int64_t V = (int64_t)d->grid * d->grid * d->grid;
    if(V > (int64_t)1 << 30)
        return (g_err = "mi_synth: grid too large", -1);
    return 0;
}
}  // namespace

extern "C" {

const char* mi_synth_last_error(void)
{
    return g_err.c_str();
}

int mi_synth_fem_sizes_of(const mi_synth_fem_desc* d, mi_synth_fem_sizes* out)
{
    if(check_desc(d) != 0)
        return -1;
    if(!out)
        return (g_err = "mi_synth: out is NULL", -1);
    int64_t N       = d->grid;
    out->block_rows = (int32_t)(N * N * N);
    out->n_tets     = 5 * (N - 1) * (N - 1) * (N - 1);
    out->n_triplets = out->block_rows + 10 * out->n_tets + 10 * d->n_contacts;
    int64_t nbx = (N + 1) / 2, nby = (N + 1) / 2, nbz = (N + 3) / 4;
    out->n_parts = (int32_t)(nbx * nby * nbz);
    return 0;
}

int mi_synth_fem_fill(const mi_synth_fem_desc* d,
                      int32_t*                 rows,
                      int32_t*                 cols,
                      double*                  vals,
                      double*                  rhs,
                      int32_t*                 tets,
                      int32_t*                 part_ids,
                      double*                  positions)
{
    try
    {
        mi_synth_fem_sizes sz;
        if(mi_synth_fem_sizes_of(d, &sz) != 0)
            return -1;
        Grid g;
        g.N     = d->grid;
        g.V     = sz.block_rows;
        g.cells = (int64_t)(g.N - 1) * (g.N - 1) * (g.N - 1);
        g.tets  = sz.n_tets;
        const int    T   = hw_threads(d->threads);
        const double mu  = d->youngs / (2.0 * (1.0 + d->poisson));
        const double lam = d->youngs * d->poisson / ((1.0 + d->poisson) * (1.0 - 2.0 * d->poisson));
        const double dt2 = d->dt * d->dt;

        std::vector<double>  X, x, mass;
        std::vector<int32_t> tv;
        build_mesh(d, g, T, X, x, tv, mass);
        if(positions)
            std::memcpy(positions, x.data(), sizeof(double) * 3 * g.V);
        if(tets)
            std::memcpy(tets, tv.data(), sizeof(int32_t) * 4 * g.tets);

        const int64_t off_fem     = g.V;
        const int64_t off_contact = g.V + 10 * g.tets;

        if(rows && cols && vals)
        {
            // kinetic blocks
            parallel_for(g.V,
                         T,
                         [&](int64_t b, int64_t e)
                         {
                             for(int64_t v = b; v < e; ++v)
                             {
                                 rows[v]   = (int32_t)v;
                                 cols[v]   = (int32_t)v;
                                 double* o = vals + 9 * v;
                                 for(int q = 0; q < 9; ++q)
                                     o[q] = 0.0;
                                 o[0] = o[4] = o[8] = mass[v];
                             }
                         });

            // elastic blocks
            parallel_for(
                g.tets,
                T,
                [&](int64_t b, int64_t e)
                {
                    for(int64_t t = b; t < e; ++t)
                    {
                        const int32_t* q = &tv[4 * t];
                        double         Dm[9], Ds[9], Dmi[9], F[9];
                        for(int c = 0; c < 3; ++c)
                            for(int r = 0; r < 3; ++r)
                            {
                                Dm[r + 3 * c] = X[3 * q[c + 1] + r] - X[3 * q[0] + r];
                                Ds[r + 3 * c] = x[3 * q[c + 1] + r] - x[3 * q[0] + r];
                            }
                        double vol = std::fabs(det3(Dm)) / 6.0;
                        inv3(Dm, Dmi);
                        for(int c = 0; c < 3; ++c)
                            for(int r = 0; r < 3; ++r)
                            {
                                double acc = 0;
                                for(int k = 0; k < 3; ++k)
                                    acc += Ds[r + 3 * k] * Dmi[k + 3 * c];
                                F[r + 3 * c] = acc;
                            }
                        // shape gradients: dN_i = row (i-1) of Dm^{-1}, dN_0 = -sum
                        double gN[4][3];
                        for(int i = 1; i < 4; ++i)
                            for(int c = 0; c < 3; ++c)
                                gN[i][c] = Dmi[(i - 1) + 3 * c];
                        for(int c = 0; c < 3; ++c)
                            gN[0][c] = -(gN[1][c] + gN[2][c] + gN[3][c]);

                        double Hp[9][9];
                        snh_hessian_psd(F, mu, lam, Hp);
                        double K[12][12];
                        double s = vol * dt2;
                        for(int i = 0; i < 4; ++i)
                            for(int j = 0; j < 4; ++j)
                                for(int a = 0; a < 3; ++a)
                                    for(int c = 0; c < 3; ++c)
                                    {
                                        double acc = 0.0;
                                        for(int bb = 0; bb < 3; ++bb)
                                            for(int dd = 0; dd < 3; ++dd)
                                                acc += gN[i][bb] * gN[j][dd] * Hp[a + 3 * bb][c + 3 * dd];
                                        K[3 * i + a][3 * j + c] = s * acc;
                                    }
                        int64_t o = off_fem + 10 * t;
                        emit_half4(q, K, rows + o, cols + o, vals + 9 * o);
                    }
                });

            // contacts
            parallel_for(
                d->n_contacts,
                T,
                [&](int64_t b, int64_t e)
                {
                    for(int64_t s = b; s < e; ++s)
                    {
                        int32_t idx[4];
                        double  gv[12], kappa;
                        contact_stencil(d, g, tv.data(), s, idx, gv, kappa);
                        double K[12][12];
                        for(int r = 0; r < 12; ++r)
                            for(int c = 0; c < 12; ++c)
                                K[r][c] = dt2 * kappa * gv[r] * gv[c];
                        int64_t o = off_contact + 10 * s;
                        emit_half4(idx, K, rows + o, cols + o, vals + 9 * o);
                    }
                });
        }

        if(rhs)
            parallel_for(3 * g.V,
                         T,
                         [&](int64_t b, int64_t e)
                         {
                             for(int64_t i = b; i < e; ++i)
                             {
                                 Rng rng(d->seed, TAG_RHS, (uint64_t)i);
                                 rhs[i] = rng.uniform(-1.0, 1.0);
                             }
                         });

        if(part_ids)
        {
            // 2x2x4 vertex blocks, numbered in Morton order of the block
            // coordinates so that 16 consecutive parts form a compact group
            // (the MAS hierarchy clusters consecutive part ids at each level)
            int64_t nbx = (g.N + 1) / 2, nby = (g.N + 1) / 2, nbz = (g.N + 3) / 4;
            std::vector<std::pair<uint32_t, int32_t>> codes;
            codes.reserve(nbx * nby * nbz);
            for(int64_t bz = 0; bz < nbz; ++bz)
                for(int64_t by = 0; by < nby; ++by)
                    for(int64_t bx = 0; bx < nbx; ++bx)
                    {
                        uint32_t m = part1by2((uint32_t)bx) | (part1by2((uint32_t)by) << 1)
                                     | (part1by2((uint32_t)bz) << 2);
                        codes.push_back({m, (int32_t)(bx + nbx * (by + nby * bz))});
                    }
            std::sort(codes.begin(), codes.end());
            std::vector<int32_t> rank(codes.size());
            for(size_t r = 0; r < codes.size(); ++r)
                rank[codes[r].second] = (int32_t)r;
            for(int64_t v = 0; v < g.V; ++v)
            {
                int i = (int)(v % g.N), j = (int)((v / g.N) % g.N), k = (int)(v / ((int64_t)g.N * g.N));
                part_ids[v] = rank[(i / 2) + nbx * ((j / 2) + nby * (k / 4))];
            }
        }
        return 0;
    }
    catch(const std::exception& e)
    {
        g_err = std::string("mi_synth: ") + e.what();
        return -1;
    }
}
int mi_synth_fem_aux(const mi_synth_fem_desc* d,
                     double*                  rest_positions,
                     double*                  mass,
                     int32_t*                 contact_idx,
                     double*                  contact_g,
                     double*                  contact_kappa)
{
    try
    {
        mi_synth_fem_sizes sz;
        if(mi_synth_fem_sizes_of(d, &sz) != 0)
            return -1;
        Grid g;
        g.N     = d->grid;
        g.V     = sz.block_rows;
        g.cells = (int64_t)(g.N - 1) * (g.N - 1) * (g.N - 1);
        g.tets  = sz.n_tets;
        const int            T = hw_threads(d->threads);
        std::vector<double>  X, x, m;
        std::vector<int32_t> tv;
        build_mesh(d, g, T, X, x, tv, m);
        if(rest_positions)
            std::memcpy(rest_positions, X.data(), sizeof(double) * 3 * g.V);
        if(mass)
            std::memcpy(mass, m.data(), sizeof(double) * g.V);
        if(contact_idx || contact_g || contact_kappa)
            parallel_for(d->n_contacts,
                         T,
                         [&](int64_t b, int64_t e)
                         {
                             for(int64_t s = b; s < e; ++s)
                             {
                                 int32_t idx[4];
                                 double  gv[12], kappa;
                                 contact_stencil(d, g, tv.data(), s, idx, gv, kappa);
                                 if(contact_idx)
                                     std::memcpy(contact_idx + 4 * s, idx, sizeof(idx));
                                 if(contact_g)
                                     std::memcpy(contact_g + 12 * s, gv, sizeof(gv));
                                 if(contact_kappa)
                                     contact_kappa[s] = kappa;
                             }
                         });
        return 0;
    }
    catch(const std::exception& e)
    {
        g_err = std::string("mi_synth: ") + e.what();
        return -1;
    }
}
}  // extern "C"
