// Cloth Hessian producer on gfx950: the strain-limiting Baraff-Witkin shell
// (finite_element/constitutions/strain_limiting_baraff_witkin_shell_2d.cu:78-148,
// strain_limiting_baraff_witkin_shell_2d.h), one thread per triangle, FP64:
//   rest     Dm = the rest triangle's edges in its own plane (rotated so that
//            its normal is +z, Dm2x2), B = Dm^-1, area A = |e01 x e02| / 2;
//   F        = [x1 - x0, x2 - x0] B (3x2), f_u = F e_1, f_v = F e_2;
//   stretch  per direction s = |f|: (s - 1)^2 + r (s - 1)^3 [s > 1], its
//            Hessian in f projected to PSD in closed form
//            (2 (s - 1)(3 r s - 3 r + 2) / (2 s) on the identity when s > 1,
//            plus c f^ f^T with c = (3 r s^2 - 3 r + 2) / s, or 2 when s <= 1);
//   shear    I6 = f_u . f_v, energy I6^2, its Hessian projected through the
//            analytic eigensystem of the 6x6 (the swap operator T, the mode
//            q0 of eigenvalue lambda0 = (I2 + sqrt(I2^2 + 12 I6^2)) / 2);
//   H        = A dt^2 dF/dx^T (stretch_k H_stretch + shear_k H_shear) dF/dx,
//            written as the 6 upper blocks of the 3-node stencil in the
//            half-block order (a = 0..2, b = a..2; (min, max) of the global
//            ids, transposed when reversed; utils/matrix_assembler.h:312-437);
// and the discrete shell bending of interior edges
// (finite_element/constitutions/discrete_shell_bending.cu:84-156): per hinge
// (V0, V1, V2, V3) with edge V1-V2, E = L0 kappa (theta - theta_bar)^2 / h_bar,
// its Gauss-Newton Hessian dt^2 E'' grad(theta) grad(theta)^T (PSD by
// construction, discrete_shell_bending_function.h:74-130) as the 10 upper
// blocks of the 4-node stencil.
#include "linsys.hpp"

namespace mi
{
namespace
{
constexpr int kThreads = 128;

int grid_shell(int64_t n)
{
    return std::max(1, std::min(ceil_div(n, kThreads), device_cu_count() * 16));
}

// the rest triangle's 2x2 edge matrix in its own plane (column-major)
__device__ void rest_dm(const double x0[3], const double x1[3], const double x2[3], double Dm[4])
{
    const double e1[3] = {x1[0] - x0[0], x1[1] - x0[1], x1[2] - x0[2]};
    const double e2[3] = {x2[0] - x0[0], x2[1] - x0[1], x2[2] - x0[2]};
    double       nrm[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
    const double ln     = sqrt(nrm[0] * nrm[0] + nrm[1] * nrm[1] + nrm[2] * nrm[2]);
    for(int i = 0; i < 3; ++i)
        nrm[i] /= ln;
    // rotation taking the normal to +z (Rodrigues about normal x z)
    double     R[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};  // row-major
    const double c  = nrm[2];
    if(c + 1.0 == 0.0)
    {
        R[0] = -1.0;
        R[4] = -1.0;
    }
    else
    {
        const double v[3] = {nrm[1], -nrm[0], 0.0};  // normal x (0, 0, 1)
        const double K[9] = {0, -v[2], v[1], v[2], 0, -v[0], -v[1], v[0], 0};
        for(int i = 0; i < 3; ++i)
            for(int j = 0; j < 3; ++j)
            {
                double kk = 0.0;
                for(int k = 0; k < 3; ++k)
                    kk += K[i * 3 + k] * K[k * 3 + j];
                R[i * 3 + j] += K[i * 3 + j] + kk / (1.0 + c);
            }
    }
    double u[3][2];
    const double* p[3] = {x0, x1, x2};
    for(int k = 0; k < 3; ++k)
        for(int r = 0; r < 2; ++r)
            u[k][r] = R[r * 3] * p[k][0] + R[r * 3 + 1] * p[k][1] + R[r * 3 + 2] * p[k][2];
    Dm[0] = u[1][0] - u[0][0];
    Dm[1] = u[1][1] - u[0][1];
    Dm[2] = u[2][0] - u[0][0];
    Dm[3] = u[2][1] - u[0][1];
}

// 6x6 PSD Hessian in vec(F) = (f_u, f_v) (row-major H6)
__device__ void bw_hessian(const double F[6], double stretch, double shear, double rate, double H6[36])
{
    const double* fu = F;
    const double* fv = F + 3;
    double        Hs[36] = {0}, Hh[36] = {0};
    const double  I5u = fmax(fu[0] * fu[0] + fu[1] * fu[1] + fu[2] * fu[2], 1e-12);
    const double  I5v = fmax(fv[0] * fv[0] + fv[1] * fv[1] + fv[2] * fv[2], 1e-12);
    const double  su = sqrt(I5u), sv = sqrt(I5v);
    if(su > 1.0)
        for(int i = 0; i < 3; ++i)
            Hs[i * 6 + i] = 2.0 * ((su - 1.0) * (3.0 * su * rate - 3.0 * rate + 2.0) / (2.0 * su));
    if(sv > 1.0)
        for(int i = 3; i < 6; ++i)
            Hs[i * 6 + i] = 2.0 * ((sv - 1.0) * (3.0 * sv * rate - 3.0 * rate + 2.0) / (2.0 * sv));
    const double cu = su > 1.0 ? (3.0 * I5u * rate - 3.0 * rate + 2.0) / su : 2.0;
    const double cv = sv > 1.0 ? (3.0 * I5v * rate - 3.0 * rate + 2.0) / sv : 2.0;
    for(int i = 0; i < 3; ++i)
        for(int j = 0; j < 3; ++j)
        {
            Hs[i * 6 + j] += cu * (fu[i] / su) * (fu[j] / su);
            Hs[(3 + i) * 6 + 3 + j] += cv * (fv[i] / sv) * (fv[j] / sv);
        }
    // shear: g = (f_v, f_u), the swap S (S vec(F) = (f_v, f_u))
    const double I6 = fu[0] * fv[0] + fu[1] * fv[1] + fu[2] * fv[2];
    const double sg = I6 >= 0.0 ? 1.0 : -1.0;
    double       g[6];
    for(int i = 0; i < 3; ++i)
    {
        g[i]     = fv[i];
        g[3 + i] = fu[i];
    }
    // |F|^2 unguarded here (I5u, I5v above are clamped away from 0)
    const double I2   = fu[0] * fu[0] + fu[1] * fu[1] + fu[2] * fu[2] + fv[0] * fv[0] + fv[1] * fv[1] + fv[2] * fv[2];
    const double lam0 = 0.5 * (I2 + sqrt(I2 * I2 + 12.0 * I6 * I6));
    double q0[6], nq = 0.0;
    for(int i = 0; i < 6; ++i)
    {
        const double Sg = g[(i + 3) % 6];  // (S g)_i
        q0[i]           = I6 * Sg + lam0 * g[i];
        nq += q0[i] * q0[i];
    }
    nq = sqrt(nq);
    for(int i = 0; i < 6; ++i)
        q0[i] /= nq;
    // T = (I + sign(I6) S) / 2; Tq = T q0
    double Tq[6], nTq = 0.0;
    for(int i = 0; i < 6; ++i)
    {
        Tq[i] = 0.5 * (q0[i] + sg * q0[(i + 3) % 6]);
        nTq += Tq[i] * Tq[i];
    }
    const double aI6 = fabs(I6);
    for(int i = 0; i < 6; ++i)
        for(int j = 0; j < 6; ++j)
        {
            const double T = 0.5 * ((i == j ? 1.0 : 0.0) + ((j == (i + 3) % 6) ? sg : 0.0));
            Hh[i * 6 + j]  = 2.0 * (aI6 * (T - Tq[i] * Tq[j] / nTq) + lam0 * q0[i] * q0[j]);
        }
    for(int i = 0; i < 36; ++i)
        H6[i] = stretch * Hs[i] + shear * Hh[i];
}

__global__ __launch_bounds__(kThreads) void bw_shell_kernel(int64_t n,
                                                           const int32_t* __restrict__ tris,
                                                           const double* __restrict__ X,
                                                           const double* __restrict__ x,
                                                           double   stretch,
                                                           double   shear,
                                                           double   rate,
                                                           double   dt2,
                                                           int32_t  vert_offset,
                                                           int32_t* rows,
                                                           int32_t* cols,
                                                           double*  vals)
{
    for(int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x)
    {
        int32_t q[3], idx[3];
        double  Xr[3][3], xc[3][3];
        for(int k = 0; k < 3; ++k)
        {
            q[k]   = tris[3 * t + k];
            idx[k] = q[k] + vert_offset;
            for(int r = 0; r < 3; ++r)
            {
                Xr[k][r] = X[3 * (int64_t)q[k] + r];
                xc[k][r] = x[3 * (int64_t)q[k] + r];
            }
        }
        double Dm[4];
        rest_dm(Xr[0], Xr[1], Xr[2], Dm);
        const double det = Dm[0] * Dm[3] - Dm[2] * Dm[1];
        const double B[4] = {Dm[3] / det, -Dm[1] / det, -Dm[2] / det, Dm[0] / det};  // column-major inverse
        double       e1[3], e2[3], cr[3];
        for(int r = 0; r < 3; ++r)
        {
            e1[r] = Xr[1][r] - Xr[0][r];
            e2[r] = Xr[2][r] - Xr[0][r];
        }
        cr[0]              = e1[1] * e2[2] - e1[2] * e2[1];
        cr[1]              = e1[2] * e2[0] - e1[0] * e2[2];
        cr[2]              = e1[0] * e2[1] - e1[1] * e2[0];
        const double area  = 0.5 * sqrt(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]);
        // F = Ds B, stored as (f_u, f_v)
        double F[6];
        for(int r = 0; r < 3; ++r)
        {
            const double d1 = xc[1][r] - xc[0][r], d2 = xc[2][r] - xc[0][r];
            F[r]            = d1 * B[0] + d2 * B[1];
            F[3 + r]        = d1 * B[2] + d2 * B[3];
        }
        double H6[36];
        bw_hessian(F, stretch, shear, rate, H6);
        // dF/dx: vec(F) row (c, r) = sum over nodes k of w[c][k] x_k[r],
        // w[c] = (-(B0c + B1c), B0c, B1c) for column c
        double w[2][3];
        for(int c = 0; c < 2; ++c)
        {
            w[c][1] = B[2 * c];
            w[c][2] = B[2 * c + 1];
            w[c][0] = -(w[c][1] + w[c][2]);
        }
        const double scale = area * dt2;
        int64_t      slot  = 6 * t;
        for(int a = 0; a < 3; ++a)
            for(int b = a; b < 3; ++b)
            {
                double K[3][3];
                for(int r = 0; r < 3; ++r)
                    for(int s = 0; s < 3; ++s)
                    {
                        double acc = 0.0;
                        for(int ci = 0; ci < 2; ++ci)
                            for(int cj = 0; cj < 2; ++cj)
                                acc += w[ci][a] * w[cj][b] * H6[(3 * ci + r) * 6 + 3 * cj + s];
                        K[r][s] = acc * scale;
                    }
                const bool swap = idx[a] > idx[b];
                rows[slot]      = swap ? idx[b] : idx[a];
                cols[slot]      = swap ? idx[a] : idx[b];
                double* v       = vals + 9 * slot;
                for(int c = 0; c < 3; ++c)
                    for(int r = 0; r < 3; ++r)
                        v[r + 3 * c] = swap ? K[c][r] : K[r][c];
                ++slot;
            }
    }
}
// ---- discrete shell bending (hinge V0, V1 = V2 edge, V3; Gauss-Newton)
__device__ __forceinline__ void sub3(const double a[3], const double b[3], double o[3])
{
    o[0] = a[0] - b[0];
    o[1] = a[1] - b[1];
    o[2] = a[2] - b[2];
}
__device__ __forceinline__ void cross3(const double a[3], const double b[3], double o[3])
{
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}
__device__ __forceinline__ double dot3(const double a[3], const double b[3])
{
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

// d theta / d(V0, V1, V2, V3) (the closed form of utils/dihedral_angle.h:36-60)
__device__ void dihedral_gradient(const double V0[3], const double V1[3], const double V2[3], const double V3[3],
                                  double g[12])
{
    double e0[3], e1[3], e2[3], e3[3], e4[3], n1[3], n2[3];
    sub3(V2, V1, e0);
    sub3(V0, V1, e1);
    sub3(V3, V1, e2);
    sub3(V0, V2, e3);
    sub3(V3, V2, e4);
    cross3(e0, e1, n1);
    cross3(e2, e0, n2);
    const double q1 = dot3(n1, n1), q2 = dot3(n2, n2), l0 = sqrt(dot3(e0, e0));
    for(int i = 0; i < 3; ++i)
    {
        g[i]     = -l0 / q1 * n1[i];
        g[3 + i] = -dot3(e0, e3) / (l0 * q1) * n1[i] - dot3(e0, e4) / (l0 * q2) * n2[i];
        g[6 + i] = dot3(e0, e1) / (l0 * q1) * n1[i] + dot3(e0, e2) / (l0 * q2) * n2[i];
        g[9 + i] = -l0 / q2 * n2[i];
    }
}

__global__ __launch_bounds__(kThreads) void bending_kernel(int64_t n,
                                                          const int32_t* __restrict__ hinges,
                                                          const double* __restrict__ X,
                                                          const double* __restrict__ x,
                                                          double   kappa,
                                                          double   dt2,
                                                          int32_t  vert_offset,
                                                          int32_t* rows,
                                                          int32_t* cols,
                                                          double*  vals)
{
    for(int64_t h = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; h < n; h += (int64_t)gridDim.x * blockDim.x)
    {
        int32_t idx[4];
        double  Xr[4][3], xc[4][3];
        for(int k = 0; k < 4; ++k)
        {
            const int32_t q = hinges[4 * h + k];
            idx[k]          = q + vert_offset;
            for(int r = 0; r < 3; ++r)
            {
                Xr[k][r] = X[3 * (int64_t)q + r];
                xc[k][r] = x[3 * (int64_t)q + r];
            }
        }
        // rest constants (discrete_shell_bending_reference.h:8-27): hinge
        // length, h = A / (3 L0) with A the two triangles' area
        double a[3], b[3], n1[3], n2[3], e[3];
        sub3(Xr[2], Xr[1], e);
        const double L0 = sqrt(dot3(e, e));
        sub3(Xr[1], Xr[0], a);
        sub3(Xr[2], Xr[0], b);
        cross3(a, b, n1);
        sub3(Xr[2], Xr[3], a);
        sub3(Xr[1], Xr[3], b);
        cross3(a, b, n2);
        const double A    = 0.5 * (sqrt(dot3(n1, n1)) + sqrt(dot3(n2, n2)));
        const double hbar = A / 3.0 / L0;
        // E = L0 kappa (theta - theta_bar)^2 / h; Gauss-Newton Hessian
        // E'' grad grad^T, E'' = 2 L0 kappa / h, times dt^2
        double g[12];
        dihedral_gradient(xc[0], xc[1], xc[2], xc[3], g);
        const double c    = 2.0 * L0 * kappa / hbar * dt2;
        int64_t      slot = 10 * h;
        for(int p = 0; p < 4; ++p)
            for(int q = p; q < 4; ++q)
            {
                const bool swap = idx[p] > idx[q];
                rows[slot]      = swap ? idx[q] : idx[p];
                cols[slot]      = swap ? idx[p] : idx[q];
                double* v       = vals + 9 * slot;
                for(int cc = 0; cc < 3; ++cc)
                    for(int r = 0; r < 3; ++r)
                    {
                        // H_ij = (c g_i) g_j with i <= j, mirrored (exactly
                        // symmetric); the block is transposed when swapped
                        const int i = 3 * p + (swap ? cc : r), j = 3 * q + (swap ? r : cc);
                        v[r + 3 * cc] = (c * g[min(i, j)]) * g[max(i, j)];
                    }
                ++slot;
            }
    }
}
}  // namespace

void produce_bending(int64_t        n,
                     const int32_t* hinges,
                     const double*  X,
                     const double*  x,
                     double         kappa,
                     double         dt,
                     int32_t        vert_offset,
                     int32_t*       rows,
                     int32_t*       cols,
                     double*        vals,
                     hipStream_t    s)
{
    if(n > 0)
        bending_kernel<<<grid_shell(n), kThreads, 0, s>>>(n, hinges, X, x, kappa, dt * dt, vert_offset, rows, cols,
                                                         vals);
    MI_HIP(hipGetLastError());
}

void produce_bw_shell(int64_t        n,
                      const int32_t* tris,
                      const double*  X,
                      const double*  x,
                      double         stretch,
                      double         shear,
                      double         rate,
                      double         dt,
                      int32_t        vert_offset,
                      int32_t*       rows,
                      int32_t*       cols,
                      double*        vals,
                      hipStream_t    s)
{
    if(n > 0)
        bw_shell_kernel<<<grid_shell(n), kThreads, 0, s>>>(n, tris, X, x, stretch, shear, rate, dt * dt, vert_offset,
                                                          rows, cols, vals);
    MI_HIP(hipGetLastError());
}
}  // namespace mi
