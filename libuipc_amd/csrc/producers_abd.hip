// Affine-body Hessian producer on gfx950: per body, the 12x12 block
//   H = M + V dt^2 PSD(d2E/dA2)      (A part only: q = (p, a_1, a_2, a_3))
// written as the reference's 4x4 grid of 3x3 triplets with the lower blocks
// zeroed (abd_linear_subsystem.cu:21-31, 80-113; AffineBodyBDF1Kinetic,
// bdf/affine_body_bdf1_kinetic.cu:38-71; OrthoPotential,
// constitutions/ortho_potential.cu:36-70).
//
//   M      the affine body's mass matrix J^T J integrated with the density:
//          m I on block (0,0), row r of block (0, r+1) holds m xbar^T (and its
//          transpose in block (r+1, 0)), S = m <xbar xbar^T> on the diagonal
//          blocks (1,1), (2,2), (3,3) (ABDJacobiDyadicMass::add_to,
//          abd_jacobi_matrix.cu:115-135).
//   E      kappa sum_{i,j} (a_i . a_j - delta_ij)^2, a_i = q[3 + 3i .. 5 + 3i]
//          (sym/ortho_potential.inl:3-20); its Hessian in 3x3 blocks
//          H_kl = 4 kappa (delta_kl sum_j a_j a_j^T + a_l a_k^T + G_kl I),
//          G_kl = a_k . a_l - delta_kl;
//   PSD    cyclic Jacobi eigen-decomposition of the 9x9, negative eigenvalues
//          set to 0 (utils/make_spd.h:24-46).
// Fixed bodies get the identity.  One thread per body, FP64.
#include "linsys.hpp"

namespace mi
{
namespace
{
constexpr int kThreads = 64;

int grid_abd(int64_t n)
{
    return std::max(1, std::min(ceil_div(n, kThreads), device_cu_count() * 16));
}

// symmetric a (row-major 9x9) -> eigenvalues on the diagonal of a,
// eigenvectors in the columns of v; cyclic Jacobi (Golub & Van Loan 8.5)
__device__ void jacobi9(double a[81], double v[81])
{
    for(int i = 0; i < 81; ++i)
        v[i] = (i % 10 == 0) ? 1.0 : 0.0;
    for(int sweep = 0; sweep < 20; ++sweep)
    {
        double off = 0.0, diag = 0.0;
        for(int p = 0; p < 9; ++p)
        {
            diag += a[p * 9 + p] * a[p * 9 + p];
            for(int q = p + 1; q < 9; ++q)
                off += a[p * 9 + q] * a[p * 9 + q];
        }
        if(off <= 1e-32 * diag || off == 0.0)
            break;
        for(int p = 0; p < 8; ++p)
            for(int q = p + 1; q < 9; ++q)
            {
                const double apq = a[p * 9 + q];
                if(apq == 0.0)
                    continue;
                const double theta = (a[q * 9 + q] - a[p * 9 + p]) / (2.0 * apq);
                const double t     = (theta >= 0.0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                const double c     = 1.0 / sqrt(t * t + 1.0);
                const double s     = t * c;
                for(int k = 0; k < 9; ++k)
                {
                    // columns p and q
                    const double akp = a[k * 9 + p], akq = a[k * 9 + q];
                    a[k * 9 + p]     = c * akp - s * akq;
                    a[k * 9 + q]     = s * akp + c * akq;
                }
                for(int k = 0; k < 9; ++k)
                {
                    // rows p and q
                    const double apk = a[p * 9 + k], aqk = a[q * 9 + k];
                    a[p * 9 + k]     = c * apk - s * aqk;
                    a[q * 9 + k]     = s * apk + c * aqk;
                }
                for(int k = 0; k < 9; ++k)
                {
                    const double vkp = v[k * 9 + p], vkq = v[k * 9 + q];
                    v[k * 9 + p]     = c * vkp - s * vkq;
                    v[k * 9 + q]     = s * vkp + c * vkq;
                }
            }
    }
}

__global__ __launch_bounds__(kThreads) void abd_kernel(int64_t n,
                                                      const double* __restrict__ q,
                                                      const double* __restrict__ dyadic,
                                                      const double* __restrict__ volume,
                                                      const double* __restrict__ kappa,
                                                      const int32_t* __restrict__ is_fixed,
                                                      double   dt2,
                                                      int32_t  body_offset,
                                                      int32_t* rows,
                                                      int32_t* cols,
                                                      double*  vals)
{
    for(int64_t I = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; I < n; I += (int64_t)gridDim.x * blockDim.x)
    {
        double     H[144];  // row-major 12x12
        const bool fixed = is_fixed && is_fixed[I] != 0;
        for(int i = 0; i < 144; ++i)
            H[i] = (fixed && i % 13 == 0) ? 1.0 : 0.0;
        if(!fixed)
        {
            // mass matrix
            const double* D = dyadic + 10 * I;
            const double  m = D[0];
            const double  S[9] = {D[4], D[5], D[6], D[5], D[7], D[8], D[6], D[8], D[9]};
            for(int r = 0; r < 3; ++r)
            {
                H[r * 12 + r] += m;
                for(int c = 0; c < 3; ++c)
                {
                    H[r * 12 + 3 + 3 * r + c] += D[1 + c];
                    H[(3 + 3 * r + c) * 12 + r] += D[1 + c];
                    for(int d = 0; d < 3; ++d)
                        H[(3 + 3 * r + c) * 12 + 3 + 3 * r + d] += S[c * 3 + d];
                }
            }
            // ortho-potential Hessian of the A part, projected to PSD
            const double* qa = q + 12 * I + 3;
            const double  k4 = 4.0 * kappa[I];
            double        G[9], AtA[9];
            for(int k = 0; k < 3; ++k)
                for(int l = 0; l < 3; ++l)
                {
                    double g = 0.0;
                    for(int e = 0; e < 3; ++e)
                        g += qa[3 * k + e] * qa[3 * l + e];
                    G[k * 3 + l] = g - (k == l ? 1.0 : 0.0);
                }
            for(int m0 = 0; m0 < 3; ++m0)
                for(int n0 = 0; n0 < 3; ++n0)
                {
                    double s = 0.0;
                    for(int j = 0; j < 3; ++j)
                        s += qa[3 * j + m0] * qa[3 * j + n0];
                    AtA[m0 * 3 + n0] = s;
                }
            double a9[81], v9[81];
            for(int k = 0; k < 3; ++k)
                for(int l = 0; l < 3; ++l)
                    for(int m0 = 0; m0 < 3; ++m0)
                        for(int n0 = 0; n0 < 3; ++n0)
                            a9[(3 * k + m0) * 9 + 3 * l + n0] =
                                k4 * ((k == l ? AtA[m0 * 3 + n0] : 0.0) + qa[3 * l + m0] * qa[3 * k + n0]
                                      + (m0 == n0 ? G[k * 3 + l] : 0.0));
            jacobi9(a9, v9);
            const double scale = volume[I] * dt2;
            double       w[9];
            for(int e = 0; e < 9; ++e)
                w[e] = a9[e * 10] < 0.0 ? 0.0 : a9[e * 10] * scale;
            for(int i = 0; i < 9; ++i)
                for(int j = i; j < 9; ++j)
                {
                    double s = 0.0;
                    for(int e = 0; e < 9; ++e)
                        s += v9[i * 9 + e] * w[e] * v9[j * 9 + e];
                    H[(3 + i) * 12 + 3 + j] += s;
                    if(j != i)
                        H[(3 + j) * 12 + 3 + i] += s;
                }
        }
        // the 4x4 grid of 3x3 blocks, row-major, lower blocks zero
        for(int ii = 0; ii < 4; ++ii)
            for(int jj = 0; jj < 4; ++jj)
            {
                const int64_t t = 16 * I + 4 * ii + jj;
                rows[t]         = body_offset + 4 * (int32_t)I + ii;
                cols[t]         = body_offset + 4 * (int32_t)I + jj;
                double* o       = vals + 9 * t;
                for(int c = 0; c < 3; ++c)
                    for(int r = 0; r < 3; ++r)
                        o[r + 3 * c] = ii > jj ? 0.0 : H[(3 * ii + r) * 12 + 3 * jj + c];
            }
    }
}
}  // namespace

void produce_abd(int64_t        n,
                 const double*  q,
                 const double*  dyadic,
                 const double*  volume,
                 const double*  kappa,
                 const int32_t* is_fixed,
                 double         dt,
                 int32_t        body_offset,
                 int32_t*       rows,
                 int32_t*       cols,
                 double*        vals,
                 hipStream_t    s)
{
    if(n > 0)
        abd_kernel<<<grid_abd(n), kThreads, 0, s>>>(n, q, dyadic, volume, kappa, is_fixed, dt * dt, body_offset,
                                                    rows, cols, vals);
    MI_HIP(hipGetLastError());
}
}  // namespace mi
