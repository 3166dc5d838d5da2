/*
 * mi_producers.h -- Hessian producers on the device: the step before the
 * linear solve, writing triplets into the caller's triplet buffer (device
 * pointers, the layout of include/mi_linsys.h) in the reference's order.
 * Declared with the rest of the C ABI in libmi_linsys.so; mi_ctx is the
 * context of mi_linsys.h.
 *
 * Every 4-node stencil writes 10 triplets, its upper blocks (a, b), a <= b,
 * in the order a = 0..3, b = a..3, at (min, max) of the two global vertex ids
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
