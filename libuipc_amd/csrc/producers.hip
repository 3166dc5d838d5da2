// Hessian producers on gfx950: the step before the solve, writing triplets
// straight into the triplet buffer in the reference's layout.
//
//   kinetic  m_i I per vertex (finite_element/bdf/finite_element_bdf1_kinetic.cu:60-70)
//   SNH      per tet, V dt^2 times the PSD projection of the stable Neo-Hookean
//            Hessian, Psi = mu/2 (tr F^T F - 3) + lambda/2 (J - alpha)^2,
//            alpha = 1 + mu / lambda (finite_element/constitutions/
//            stable_neo_hookean_3d.cu:100-264), projected through its analytic
//            eigensystem: in the SVD frame F = U S V^T the Hessian splits into
//            three scaling modes (eigenvectors of a 3x3 matrix in the singular
//            values) and, per pair (i, j) with k the third index, a twist mode
//            (antisymmetric, mu + lambda (J - alpha) s_k) and a flip mode
//            (symmetric, mu - lambda (J - alpha) s_k); negative eigenvalues
//            are clamped to 0
//   contact  dt^2 kappa g g^T per vertex-triangle stencil (the synthetic
//            stand-in for contact_system/simplex_normal_contact.cu:74-161)
// and, for the line search after the solve, the kinetic and SNH energies
// (each reduced in a fixed order: per-workgroup partials, then one workgroup)
// and the trial step x = x0 + alpha dx.
//
// Each 4-node stencil writes its 10 upper blocks as the reference's half-block
// writer does (utils/matrix_assembler.h:312-437): slot (a, b), a <= b, in the
// order a = 0..3, b = a..3; the block is written at (min, max) of the two
// global vertex ids, transposed when the stencil order is reversed.
// One thread per tet / stencil / vertex; all arithmetic in FP64.
#include "linsys.hpp"

namespace mi
{
namespace
{
constexpr int kThreads = 256;

int grid_of(int64_t n)
{
    return std::max(1, std::min(ceil_div(n, kThreads), device_cu_count() * 16));
}

__device__ __forceinline__ double det3(const double m[9])  // column-major
{
    return m[0] * (m[4] * m[8] - m[7] * m[5]) - m[3] * (m[1] * m[8] - m[7] * m[2])
           + m[6] * (m[1] * m[5] - m[4] * m[2]);
}

__device__ __forceinline__ void inv3(const double m[9], double o[9])
{
    const double id = 1.0 / det3(m);
    o[0]            = (m[4] * m[8] - m[7] * m[5]) * id;
    o[3]            = -(m[3] * m[8] - m[6] * m[5]) * id;
    o[6]            = (m[3] * m[7] - m[6] * m[4]) * id;
    o[1]            = -(m[1] * m[8] - m[7] * m[2]) * id;
    o[4]            = (m[0] * m[8] - m[6] * m[2]) * id;
    o[7]            = -(m[0] * m[7] - m[6] * m[1]) * id;
    o[2]            = (m[1] * m[5] - m[4] * m[2]) * id;
    o[5]            = -(m[0] * m[5] - m[3] * m[2]) * id;
    o[8]            = (m[0] * m[4] - m[3] * m[1]) * id;
}

// cyclic Jacobi eigen-decomposition of a symmetric 3x3 matrix (row-major a):
// a -> diag(w), eigenvectors in the columns of q (row-major)
__device__ void jacobi3(double a[3][3], double w[3], double q[3][3])
{
    for(int i = 0; i < 3; ++i)
        for(int j = 0; j < 3; ++j)
            q[i][j] = i == j ? 1.0 : 0.0;
    for(int sweep = 0; sweep < 20; ++sweep)
    {
        const double off = a[0][1] * a[0][1] + a[0][2] * a[0][2] + a[1][2] * a[1][2];
        const double dia = a[0][0] * a[0][0] + a[1][1] * a[1][1] + a[2][2] * a[2][2];
        if(off <= 1e-32 * dia || off == 0.0)
            break;
        for(int p = 0; p < 2; ++p)
            for(int r = p + 1; r < 3; ++r)
            {
                const double apr = a[p][r];
                if(apr == 0.0)
                    continue;
                const double theta = (a[r][r] - a[p][p]) / (2.0 * apr);
                const double t     = (theta >= 0.0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                const double c     = 1.0 / sqrt(t * t + 1.0);
                const double s     = t * c;
                for(int k = 0; k < 3; ++k)
                {
                    const double akp = a[k][p], akr = a[k][r];
                    a[k][p]          = c * akp - s * akr;
                    a[k][r]          = s * akp + c * akr;
                }
                for(int k = 0; k < 3; ++k)
                {
                    const double apk = a[p][k], ark = a[r][k];
                    a[p][k]          = c * apk - s * ark;
                    a[r][k]          = s * apk + c * ark;
                }
                for(int k = 0; k < 3; ++k)
                {
                    const double qkp = q[k][p], qkr = q[k][r];
                    q[k][p]          = c * qkp - s * qkr;
                    q[k][r]          = s * qkp + c * qkr;
                }
            }
    }
    for(int i = 0; i < 3; ++i)
        w[i] = a[i][i];
}

// F = U diag(S) V^T with U, V rotations (S[2] carries the sign of det F);
// F, U, V column-major
__device__ void svd3(const double F[9], double U[9], double S[3], double V[9])
{
    double C[3][3], w[3], q[3][3];
    for(int i = 0; i < 3; ++i)
        for(int j = 0; j < 3; ++j)
            C[i][j] = F[3 * i] * F[3 * j] + F[3 * i + 1] * F[3 * j + 1] + F[3 * i + 2] * F[3 * j + 2];
    jacobi3(C, w, q);
    // descending singular values
    int o[3] = {0, 1, 2};
    for(int i = 0; i < 2; ++i)
        for(int j = i + 1; j < 3; ++j)
            if(w[o[j]] > w[o[i]])
            {
                const int t = o[i];
                o[i]        = o[j];
                o[j]        = t;
            }
    for(int c = 0; c < 3; ++c)
        for(int r = 0; r < 3; ++r)
            V[r + 3 * c] = q[r][o[c]];
    // V a rotation
    const double dv = det3(V);
    if(dv < 0.0)
        for(int r = 0; r < 3; ++r)
            V[r + 6] = -V[r + 6];
    // u_c = F v_c / s_c for the two largest; u_2 = u_0 x u_1
    double Fv[3][3];
    for(int c = 0; c < 3; ++c)
        for(int r = 0; r < 3; ++r)
            Fv[c][r] = F[r] * V[3 * c] + F[r + 3] * V[3 * c + 1] + F[r + 6] * V[3 * c + 2];
    double n0 = sqrt(Fv[0][0] * Fv[0][0] + Fv[0][1] * Fv[0][1] + Fv[0][2] * Fv[0][2]);
    for(int r = 0; r < 3; ++r)
        U[r] = Fv[0][r] / n0;
    double d01 = U[0] * Fv[1][0] + U[1] * Fv[1][1] + U[2] * Fv[1][2];
    double u1[3];
    for(int r = 0; r < 3; ++r)
        u1[r] = Fv[1][r] - d01 * U[r];
    double n1 = sqrt(u1[0] * u1[0] + u1[1] * u1[1] + u1[2] * u1[2]);
    for(int r = 0; r < 3; ++r)
        U[3 + r] = u1[r] / n1;
    U[6] = U[1] * U[5] - U[2] * U[4];
    U[7] = U[2] * U[3] - U[0] * U[5];
    U[8] = U[0] * U[4] - U[1] * U[3];
    S[0] = n0;
    S[1] = U[3] * Fv[1][0] + U[4] * Fv[1][1] + U[5] * Fv[1][2];
    S[2] = U[6] * Fv[2][0] + U[7] * Fv[2][1] + U[8] * Fv[2][2];
}

// writes slot (a, b) of a 4-node stencil: (min, max) of the global ids, the
// block transposed when the stencil order is reversed; K_ab is row-major
__device__ __forceinline__ void emit_block(const int32_t idx[4],
                                           int           a,
                                           int           b,
                                           const double  K[3][3],
                                           int64_t       t,
                                           int32_t*      rows,
                                           int32_t*      cols,
                                           double*       vals)
{
    const bool swap = idx[a] > idx[b];
    rows[t]         = swap ? idx[b] : idx[a];
    cols[t]         = swap ? idx[a] : idx[b];
    double* v       = vals + 9 * t;
#pragma unroll
    for(int c = 0; c < 3; ++c)
#pragma unroll
        for(int r = 0; r < 3; ++r)
            v[r + 3 * c] = swap ? K[c][r] : K[r][c];
}

__global__ __launch_bounds__(kThreads) void kinetic_kernel(int32_t V,
                                                          const double* __restrict__ mass,
                                                          int32_t  vert_offset,
                                                          int32_t* rows,
                                                          int32_t* cols,
                                                          double*  vals)
{
    for(int32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < V; v += gridDim.x * blockDim.x)
    {
        rows[v]   = v + vert_offset;
        cols[v]   = v + vert_offset;
        double* o = vals + 9 * (int64_t)v;
#pragma unroll
        for(int q = 0; q < 9; ++q)
            o[q] = (q % 4 == 0) ? mass[v] : 0.0;
    }
}

__global__ __launch_bounds__(kThreads) void snh_kernel(int64_t n_tets,
                                                      const int32_t* __restrict__ tets,
                                                      const double* __restrict__ X,
                                                      const double* __restrict__ x,
                                                      double   mu,
                                                      double   lambda,
                                                      double   dt2,
                                                      int32_t  vert_offset,
                                                      int32_t* rows,
                                                      int32_t* cols,
                                                      double*  vals)
{
    for(int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n_tets; t += (int64_t)gridDim.x * blockDim.x)
    {
        int32_t q[4], idx[4];
#pragma unroll
        for(int c = 0; c < 4; ++c)
        {
            q[c]   = tets[4 * t + c];
            idx[c] = q[c] + vert_offset;
        }
        double Dm[9], Ds[9], Dmi[9], F[9];
#pragma unroll
        for(int c = 0; c < 3; ++c)
#pragma unroll
            for(int r = 0; r < 3; ++r)
            {
                Dm[r + 3 * c] = X[3 * (int64_t)q[c + 1] + r] - X[3 * (int64_t)q[0] + r];
                Ds[r + 3 * c] = x[3 * (int64_t)q[c + 1] + r] - x[3 * (int64_t)q[0] + r];
            }
        const double vol = fabs(det3(Dm)) / 6.0;
        inv3(Dm, Dmi);
#pragma unroll
        for(int c = 0; c < 3; ++c)
#pragma unroll
            for(int r = 0; r < 3; ++r)
                F[r + 3 * c] = Ds[r] * Dmi[3 * c] + Ds[r + 3] * Dmi[1 + 3 * c] + Ds[r + 6] * Dmi[2 + 3 * c];
        // shape gradients: dN_i = row (i - 1) of Dm^-1, dN_0 = -(dN_1 + dN_2 + dN_3)
        double gN[4][3];
#pragma unroll
        for(int i = 1; i < 4; ++i)
#pragma unroll
            for(int c = 0; c < 3; ++c)
                gN[i][c] = Dmi[(i - 1) + 3 * c];
#pragma unroll
        for(int c = 0; c < 3; ++c)
            gN[0][c] = -(gN[1][c] + gN[2][c] + gN[3][c]);

        double U[9], S[3], Vm[9];
        svd3(F, U, S, Vm);
        const double J     = S[0] * S[1] * S[2];
        const double alpha = 1.0 + mu / lambda;
        const double s     = lambda * (J - alpha);
        const double scale = vol * dt2;
        // sa[i][k] = V_k . dN_i
        double sa[4][3];
#pragma unroll
        for(int i = 0; i < 4; ++i)
#pragma unroll
            for(int k = 0; k < 3; ++k)
                sa[i][k] = Vm[3 * k] * gN[i][0] + Vm[3 * k + 1] * gN[i][1] + Vm[3 * k + 2] * gN[i][2];
        // scaling modes: the 3x3 Hessian of Psi in the singular values
        double A[3][3], ew[3], ev[3][3];
        A[0][0] = mu + lambda * (S[1] * S[2]) * (S[1] * S[2]);
        A[1][1] = mu + lambda * (S[0] * S[2]) * (S[0] * S[2]);
        A[2][2] = mu + lambda * (S[0] * S[1]) * (S[0] * S[1]);
        const double c2 = lambda * (2.0 * J - alpha);
        A[0][1] = A[1][0] = c2 * S[2];
        A[0][2] = A[2][0] = c2 * S[1];
        A[1][2] = A[2][1] = c2 * S[0];
        jacobi3(A, ew, ev);
        // the 10 upper blocks, each summed over the 9 modes: lam w_a w_b^T
        constexpr int kPairs[3][2] = {{1, 2}, {0, 2}, {0, 1}};  // (i, j) of the mode with third index k
        const double  r2           = 0.70710678118654752440;  // 1 / sqrt(2)
        int64_t       slot         = 10 * t;
#pragma unroll
        for(int a = 0; a < 4; ++a)
#pragma unroll
            for(int b = a; b < 4; ++b)
            {
                double K[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
                // scaling modes: w_i = U (e_m (.) sa_i)
#pragma unroll
                for(int m = 0; m < 3; ++m)
                {
                    const double l = fmax(ew[m], 0.0) * scale;
                    double       wa[3], wb[3];
#pragma unroll
                    for(int r = 0; r < 3; ++r)
                    {
                        wa[r] = U[r] * ev[0][m] * sa[a][0] + U[r + 3] * ev[1][m] * sa[a][1]
                                + U[r + 6] * ev[2][m] * sa[a][2];
                        wb[r] = U[r] * ev[0][m] * sa[b][0] + U[r + 3] * ev[1][m] * sa[b][1]
                                + U[r + 6] * ev[2][m] * sa[b][2];
                    }
#pragma unroll
                    for(int r = 0; r < 3; ++r)
#pragma unroll
                        for(int c = 0; c < 3; ++c)
                            K[r][c] += l * wa[r] * wb[c];
                }
                // twist (-) and flip (+) modes of the pair (i, j), third index k:
                // w = (U_i sa[j] -/+ U_j sa[i]) / sqrt(2)
#pragma unroll
                for(int k = 0; k < 3; ++k)
                {
                    const int i = kPairs[k][0], j = kPairs[k][1];
#pragma unroll
                    for(int flip = 0; flip < 2; ++flip)
                    {
                        const double l  = fmax(flip ? mu - s * S[k] : mu + s * S[k], 0.0) * scale;
                        const double sg = flip ? 1.0 : -1.0;
                        double       wa[3], wb[3];
#pragma unroll
                        for(int r = 0; r < 3; ++r)
                        {
                            wa[r] = r2 * (U[r + 3 * i] * sa[a][j] + sg * U[r + 3 * j] * sa[a][i]);
                            wb[r] = r2 * (U[r + 3 * i] * sa[b][j] + sg * U[r + 3 * j] * sa[b][i]);
                        }
#pragma unroll
                        for(int r = 0; r < 3; ++r)
#pragma unroll
                            for(int c = 0; c < 3; ++c)
                                K[r][c] += l * wa[r] * wb[c];
                    }
                }
                emit_block(idx, a, b, K, slot++, rows, cols, vals);
            }
    }
}

__global__ __launch_bounds__(kThreads) void contact_kernel(int64_t n,
                                                          const int32_t* __restrict__ cidx,
                                                          const double* __restrict__ g,
                                                          const double* __restrict__ kappa,
                                                          double   dt2,
                                                          int32_t  vert_offset,
                                                          int32_t* rows,
                                                          int32_t* cols,
                                                          double*  vals)
{
    for(int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < n; s += (int64_t)gridDim.x * blockDim.x)
    {
        int32_t idx[4];
#pragma unroll
        for(int c = 0; c < 4; ++c)
            idx[c] = cidx[4 * s + c] + vert_offset;
        double gv[12];
#pragma unroll
        for(int c = 0; c < 12; ++c)
            gv[c] = g[12 * s + c];
        const double k    = dt2 * kappa[s];
        int64_t      slot = 10 * s;
#pragma unroll
        for(int a = 0; a < 4; ++a)
#pragma unroll
            for(int b = a; b < 4; ++b)
            {
                // the block in its written orientation (L, R) = (a, b) or (b, a),
                // each entry k g_(L, r) g_(R, c) as one product chain
                const bool swap = idx[a] > idx[b];
                const int  L = swap ? b : a, R = swap ? a : b;
                rows[slot]   = idx[L];
                cols[slot]   = idx[R];
                double* v    = vals + 9 * slot;
#pragma unroll
                for(int c = 0; c < 3; ++c)
#pragma unroll
                    for(int r = 0; r < 3; ++r)
                        v[r + 3 * c] = k * gv[3 * L + r] * gv[3 * R + c];
                ++slot;
            }
    }
}
// ---- line-search side: energies and the trial step
constexpr int kEnergyBlocks = 1024;

int energy_grid(int64_t n)
{
    return (int)std::max<int64_t>(1, std::min<int64_t>(ceil_div(n, kThreads), kEnergyBlocks));
}

// FiniteElementBDF1Kinetic_do_compute_energy_kernel: 1/2 m |x - x_tilde|^2
__global__ __launch_bounds__(kThreads) void kinetic_energy_kernel(int32_t V,
                                                                 const double* __restrict__ x,
                                                                 const double* __restrict__ xt,
                                                                 const double* __restrict__ m,
                                                                 double* part)
{
    __shared__ double lds[kThreads / kWave];
    double            acc = 0.0;
    for(int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; v < V; v += (int64_t)gridDim.x * blockDim.x)
    {
        const double d0 = x[3 * v] - xt[3 * v], d1 = x[3 * v + 1] - xt[3 * v + 1], d2 = x[3 * v + 2] - xt[3 * v + 2];
        acc += 0.5 * m[v] * (d0 * d0 + d1 * d1 + d2 * d2);
    }
    acc = block_sum<kThreads>(acc, lds);
    if(threadIdx.x == 0)
        part[blockIdx.x] = acc;
}

// StableNeoHookean3D_do_compute_energy_kernel: dt^2 V_rest Psi(F),
// Psi = 1/2 (mu (|F|^2 - 3) + lambda (J - 1 - mu / lambda)^2), J = det F
__global__ __launch_bounds__(kThreads) void snh_energy_kernel(int64_t n_tets,
                                                             const int32_t* __restrict__ tets,
                                                             const double* __restrict__ X,
                                                             const double* __restrict__ x,
                                                             double  mu,
                                                             double  lambda,
                                                             double  dt2,
                                                             double* part)
{
    __shared__ double lds[kThreads / kWave];
    double            acc = 0.0;
    for(int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n_tets; t += (int64_t)gridDim.x * blockDim.x)
    {
        int32_t q[4];
#pragma unroll
        for(int c = 0; c < 4; ++c)
            q[c] = tets[4 * t + c];
        double Dm[9], Ds[9], Dmi[9], F[9];
#pragma unroll
        for(int c = 0; c < 3; ++c)
#pragma unroll
            for(int r = 0; r < 3; ++r)
            {
                Dm[r + 3 * c] = X[3 * (int64_t)q[c + 1] + r] - X[3 * (int64_t)q[0] + r];
                Ds[r + 3 * c] = x[3 * (int64_t)q[c + 1] + r] - x[3 * (int64_t)q[0] + r];
            }
        const double vol = fabs(det3(Dm)) / 6.0;
        inv3(Dm, Dmi);
        double Ic = 0.0;
#pragma unroll
        for(int c = 0; c < 3; ++c)
#pragma unroll
            for(int r = 0; r < 3; ++r)
            {
                F[r + 3 * c] = Ds[r] * Dmi[3 * c] + Ds[r + 3] * Dmi[1 + 3 * c] + Ds[r + 6] * Dmi[2 + 3 * c];
                Ic += F[r + 3 * c] * F[r + 3 * c];
            }
        const double Jm = det3(F) - 1.0 - mu / lambda;
        acc += 0.5 * (mu * (Ic - 3.0) + lambda * Jm * Jm) * dt2 * vol;
    }
    acc = block_sum<kThreads>(acc, lds);
    if(threadIdx.x == 0)
        part[blockIdx.x] = acc;
}

// the partials in a fixed order (one workgroup), so the total is reproducible
__global__ __launch_bounds__(kThreads) void energy_total_kernel(const double* __restrict__ part, int n, double* out)
{
    __shared__ double lds[kThreads / kWave];
    double            acc = 0.0;
    for(int i = threadIdx.x; i < n; i += kThreads)
        acc += part[i];
    acc = block_sum<kThreads>(acc, lds);
    if(threadIdx.x == 0)
        *out = acc;
}

__global__ __launch_bounds__(kThreads) void step_forward_kernel(const double* __restrict__ x0,
                                                               const double* __restrict__ dx,
                                                               double  alpha,
                                                               double* x,
                                                               int64_t n)
{
    for(int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        x[i] = x0[i] + alpha * dx[i];
}
}  // namespace

void energy_kinetic(int32_t V, const double* x, const double* xt, const double* m, double* energy,
                    DBuf<double>& part, hipStream_t s)
{
    const int g = energy_grid(V);
    part.resize(kEnergyBlocks);
    kinetic_energy_kernel<<<g, kThreads, 0, s>>>(V, x, xt, m, part.data());
    energy_total_kernel<<<1, kThreads, 0, s>>>(part.data(), g, energy);
    MI_HIP(hipGetLastError());
}

void energy_snh(int64_t        n_tets,
                const int32_t* tets,
                const double*  X,
                const double*  x,
                double         mu,
                double         lambda,
                double         dt,
                double*        energy,
                DBuf<double>&  part,
                hipStream_t    s)
{
    const int g = energy_grid(n_tets);
    part.resize(kEnergyBlocks);
    snh_energy_kernel<<<g, kThreads, 0, s>>>(n_tets, tets, X, x, mu, lambda, dt * dt, part.data());
    energy_total_kernel<<<1, kThreads, 0, s>>>(part.data(), g, energy);
    MI_HIP(hipGetLastError());
}

void step_forward(const double* x0, const double* dx, double alpha, double* x, int64_t n, hipStream_t s)
{
    if(n > 0)
        step_forward_kernel<<<grid_of(n), kThreads, 0, s>>>(x0, dx, alpha, x, n);
    MI_HIP(hipGetLastError());
}

void produce_kinetic(int32_t V, const double* mass, int32_t vert_offset, int32_t* rows, int32_t* cols, double* vals,
                     hipStream_t s)
{
    if(V > 0)
        kinetic_kernel<<<grid_of(V), kThreads, 0, s>>>(V, mass, vert_offset, rows, cols, vals);
    MI_HIP(hipGetLastError());
}

void produce_snh(int64_t        n_tets,
                 const int32_t* tets,
                 const double*  X,
                 const double*  x,
                 double         mu,
                 double         lambda,
                 double         dt,
                 int32_t        vert_offset,
                 int32_t*       rows,
                 int32_t*       cols,
                 double*        vals,
                 hipStream_t    s)
{
    if(n_tets > 0)
        snh_kernel<<<grid_of(n_tets), kThreads, 0, s>>>(n_tets, tets, X, x, mu, lambda, dt * dt, vert_offset, rows,
                                                        cols, vals);
    MI_HIP(hipGetLastError());
}

void produce_contact(int64_t        n,
                     const int32_t* stencils,
                     const double*  g,
                     const double*  kappa,
                     double         dt,
                     int32_t        vert_offset,
                     int32_t*       rows,
                     int32_t*       cols,
                     double*        vals,
                     hipStream_t    s)
{
    if(n > 0)
        contact_kernel<<<grid_of(n), kThreads, 0, s>>>(n, stencils, g, kappa, dt * dt, vert_offset, rows, cols, vals);
    MI_HIP(hipGetLastError());
}
}  // namespace mi
