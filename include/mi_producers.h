/*
 * mi_producers.h -- Hessian producers on the device: the step before the
 * linear solve, writing triplets into the caller's triplet buffer (device
 * pointers, the layout of include/mi_linsys.h) in the reference's order.
 * Declared with the rest of the C ABI in libmi_linsys.so; mi_ctx is the
 * context of mi_linsys.h.
 *
 * Every N-node stencil writes N (N + 1) / 2 triplets, its upper blocks (a, b),
 * a <= b, in the order a = 0..N-1, b = a..N-1, at (min, max) of the two global vertex ids
 * and transposed when the stencil order is reversed: the half-block writer
 * of utils/matrix_assembler.h:312-437 (TripletMatrixAssembler::ProxyRangeHalf).
 * Vertex ids in the inputs are local to the subsystem; vert_offset (its first
 * block row) is added on output.
 */
#ifndef MI_PRODUCERS_H
#define MI_PRODUCERS_H

#include <stdint.h>

#include "mi_linsys.h"

#ifdef __cplusplus
extern "C" {
#endif

/* The BDF1 kinetic Hessian: block (v, v) = mass[v] I for v < vert_num
 * (finite_element/bdf/finite_element_bdf1_kinetic.cu:60-70), one triplet per
 * vertex at rows[v], cols[v], vals[9 v]. */
int mi_produce_kinetic(mi_ctx*       ctx,
                       int32_t       vert_num,
                       const double* mass,
                       int32_t       vert_offset,
                       int32_t*      rows,
                       int32_t*      cols,
                       double*       vals,
                       void*         stream);

/* Stable Neo-Hookean tets (finite_element/constitutions/stable_neo_hookean_3d.cu:100-264):
 * per tet t (tets[4 t .. 4 t + 3], rest volume > 0 with vertex 0 as origin),
 * F = Ds Dm^-1 from rest_positions and positions (3 doubles per vertex), and
 * the 12x12 block V dt^2 d2Psi/dx2 with the Hessian projected to PSD through
 * its analytic eigensystem (clamped eigenvalues), Psi = mu/2 (tr F^T F - 3)
 * + lambda/2 (J - 1 - mu/lambda)^2.  Writes triplets 10 t .. 10 t + 9. */
int mi_produce_snh(mi_ctx*        ctx,
                   int64_t        n_tets,
                   const int32_t* tets,
                   const double*  rest_positions,
                   const double*  positions,
                   double         mu,
                   double         lambda,
                   double         dt,
                   int32_t        vert_offset,
                   int32_t*       rows,
                   int32_t*       cols,
                   double*        vals,
                   void*          stream);

/* Vertex-triangle contact stencils: per stencil s (stencils[4 s .. 4 s + 3],
 * the vertex then the triangle), the 12x12 block dt^2 kappa[s] g g^T with
 * g = g[12 s .. 12 s + 11]: the synthetic stand-in for the barrier Hessians
 * of contact_system/simplex_normal_contact.cu:74-161 (same stencil and
 * half_block<4> layout).  Writes triplets 10 s .. 10 s + 9. */
int mi_produce_contact(mi_ctx*        ctx,
                       int64_t        n_stencils,
                       const int32_t* stencils,
                       const double*  g,
                       const double*  kappa,
                       double         dt,
                       int32_t        vert_offset,
                       int32_t*       rows,
                       int32_t*       cols,
                       double*        vals,
                       void*          stream);

/* Affine bodies (12 DoFs q = (p, a_1, a_2, a_3), a_i the rows of A, x = p + A xbar):
 * per body I, the 12x12 block H = M + volume dt^2 PSD(d2E/dA2) written as 16
 * triplets 16 I .. 16 I + 15, a row-major 4x4 grid of 3x3 blocks at block rows
 * body_offset + 4 I + (0..3), the blocks below the diagonal set to zero
 * (abd_linear_subsystem.cu:21-31, 80-113).  M is the affine mass matrix of
 * AffineBodyBDF1Kinetic (affine_body/bdf/affine_body_bdf1_kinetic.cu:38-71,
 * ABDJacobiDyadicMass::add_to, abd_jacobi_matrix.cu:115-135) from
 * dyadic_mass[10 I ..] = (m, m xbar (3), m <xbar xbar^T> as xx, xy, xz, yy,
 * yz, zz); E = kappa sum_ij (a_i . a_j - delta_ij)^2 is the orthogonality
 * potential (affine_body/constitutions/ortho_potential.cu:36-70), its 9x9
 * Hessian projected to PSD by eigenvalue clamping (utils/make_spd.h).
 * is_fixed (may be NULL): a nonzero entry writes the identity instead. */
int mi_produce_abd(mi_ctx*        ctx,
                   int64_t        n_bodies,
                   const double*  q,
                   const double*  dyadic_mass,
                   const double*  volume,
                   const double*  kappa,
                   const int32_t* is_fixed,
                   double         dt,
                   int32_t        body_offset,
                   int32_t*       rows,
                   int32_t*       cols,
                   double*        vals,
                   void*          stream);

/* Cloth triangles, the strain-limiting Baraff-Witkin shell
 * (finite_element/constitutions/strain_limiting_baraff_witkin_shell_2d.cu:78-148):
 * per triangle t (tris[3 t .. 3 t + 2]), F = [x1 - x0, x2 - x0] Dm^-1 with Dm the
 * rest triangle's edges in its own plane, and the 9x9 block
 * A dt^2 dF/dx^T (stretch_stiffness H_stretch + shear_stiffness H_shear) dF/dx
 * (A the rest area; H_stretch, H_shear the closed-form PSD Hessians of
 * (s - 1)^2 + strain_rate (s - 1)^3 [s > 1] per direction and of I6^2,
 * strain_limiting_baraff_witkin_shell_2d.h:165-248).  Writes the 6 upper
 * blocks of the 3-node stencil, triplets 6 t .. 6 t + 5, in the half-block
 * order described above. */
int mi_produce_bw_shell(mi_ctx*        ctx,
                        int64_t        n_tris,
                        const int32_t* tris,
                        const double*  rest_positions,
                        const double*  positions,
                        double         stretch_stiffness,
                        double         shear_stiffness,
                        double         strain_rate,
                        double         dt,
                        int32_t        vert_offset,
                        int32_t*       rows,
                        int32_t*       cols,
                        double*        vals,
                        void*          stream);

/* Cloth bending, the discrete shell hinge (finite_element/constitutions/
 * discrete_shell_bending.cu:84-156): per hinge h (hinges[4 h .. 4 h + 3] =
 * V0, V1, V2, V3, the edge V1-V2 shared by triangles (V0, V1, V2) and
 * (V3, V2, V1)), E = L0 kappa (theta - theta_bar)^2 / h_bar with the rest
 * constants of discrete_shell_bending_reference.h:8-27 (L0 the rest edge
 * length, h_bar = A / (3 L0), A the two rest triangles' area) and the 12x12
 * Gauss-Newton Hessian dt^2 (2 L0 kappa / h_bar) grad(theta) grad(theta)^T
 * (discrete_shell_bending_function.h:74-130; grad(theta) as
 * utils/dihedral_angle.h:36-60).  Writes triplets 10 h .. 10 h + 9. */
int mi_produce_bending(mi_ctx*        ctx,
                       int64_t        n_hinges,
                       const int32_t* hinges,
                       const double*  rest_positions,
                       const double*  positions,
                       double         bending_stiffness,
                       double         dt,
                       int32_t        vert_offset,
                       int32_t*       rows,
                       int32_t*       cols,
                       double*        vals,
                       void*          stream);

/* ------------------------------------------------ line search, after the solve */
/* LineSearchReporter::step_forward of a vertex set: x = x0 + alpha dx over n
 * doubles (the FEM reporter's trial positions; engine/advance_ipc.cu:490-560
 * calls it for every alpha it tries). */
int mi_step_forward(mi_ctx* ctx, const double* x0, const double* dx, double alpha, double* x, int64_t n,
                    void* stream);

/* Line-search energies, LineSearcher::compute_energy
 * (line_search/line_searcher.cu:60-124): each writes its total to *energy
 * (one device double).  The sum runs over fixed per-workgroup partials and
 * then one workgroup, so a given input always gives the same bits.  Calls on
 * one context share its partials buffer: serialise them.
 *   kinetic  sum_v 1/2 mass[v] |x_v - x_tilde_v|^2
 *            (FiniteElementBDF1Kinetic_do_compute_energy_kernel,
 *             finite_element/bdf/finite_element_bdf1_kinetic.cu:8-33)
 *   SNH      sum_t dt^2 V_t Psi(F_t), V_t and F_t as mi_produce_snh,
 *            Psi = 1/2 (mu (|F|^2 - 3) + lambda (J - 1 - mu/lambda)^2)
 *            (StableNeoHookean3D_do_compute_energy_kernel,
 *             finite_element/constitutions/stable_neo_hookean_3d.cu:22-52;
 *             constitutions/detail/stable_neo_hookean_3d_snk1.inl:14-20) */
int mi_energy_kinetic(mi_ctx*       ctx,
                      int32_t       vert_num,
                      const double* x,
                      const double* x_tilde,
                      const double* mass,
                      double*       energy,
                      void*         stream);
int mi_energy_snh(mi_ctx*        ctx,
                  int64_t        n_tets,
                  const int32_t* tets,
                  const double*  rest_positions,
                  const double*  positions,
                  double         mu,
                  double         lambda,
                  double         dt,
                  double*        energy,
                  void*          stream);

#ifdef __cplusplus
}
#endif
#endif
