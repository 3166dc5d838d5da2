/*
 * mi_synth.h -- seeded synthetic linear systems in libuipc's triplet format.
 *
 * This is the producer side of the linear solve, standing in for libuipc's
 * Hessian producers (FEM constitutions, kinetic term, contact): it writes the
 * triplets (row, col, 3x3 column-major value) in the order and with the
 * upper-triangle half-block semantics the reference's producers use
 * (utils/matrix_assembler.h:312-410, fem_linear_subsystem.cu:276-355).
 * It has no counterpart in the reference's C ABI.
 *
 * All buffers are HOST memory owned by the caller; call the *_sizes function
 * first, allocate, then the *_fill function.
 */
#ifndef MI_SYNTH_H
#define MI_SYNTH_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mi_synth_fem_desc
{
    int32_t  grid;          /* vertices per axis of the unit cube (config: 16 or 64)        */
    double   jitter;        /* uniform vertex jitter, fraction of a cell (0.05)             */
    uint64_t seed;
    double   youngs;        /* E [Pa] (1e5)                                                 */
    double   poisson;       /* nu (0.4)                                                     */
    double   density;       /* rho [kg/m^3] (1e3)                                           */
    double   dt;            /* time step [s] (0.01)                                         */
    int64_t  n_contacts;    /* vertex-triangle contact stencils (0 or 2e6)                  */
    int32_t  contact_cells; /* the triangle lies in a cell within this many cells of p (2)  */
    double   kappa_lo;      /* contact stiffness, log-uniform in [kappa_lo, kappa_hi]       */
    double   kappa_hi;
    int32_t  threads;       /* host threads, 0 = all                                        */
} mi_synth_fem_desc;

typedef struct mi_synth_fem_sizes
{
    int32_t block_rows;  /* vertices                               */
    int64_t n_tets;
    int64_t n_triplets;  /* B (kinetic) + 10 n_tets + 10 n_contacts */
    int32_t n_parts;     /* geometric partition: 2x2x4 vertex blocks */
} mi_synth_fem_sizes;

int mi_synth_fem_sizes_of(const mi_synth_fem_desc* desc, mi_synth_fem_sizes* out);

/* Fills (any pointer may be NULL to skip it):
 *   rows, cols [n_triplets], vals [9 n_triplets]   the system A = M + dt^2 (K_fem + K_contact)
 *   rhs [3 block_rows]                             b ~ uniform[-1, 1]
 *   tets [4 n_tets], part_ids [block_rows]         topology and partition for MAS
 *   positions [3 block_rows]                       deformed vertex positions            */
int mi_synth_fem_fill(const mi_synth_fem_desc* desc,
                      int32_t*                 rows,
                      int32_t*                 cols,
                      double*                  vals,
                      double*                  rhs,
                      int32_t*                 tets,
                      int32_t*                 part_ids,
                      double*                  positions);

/* The inputs of the device producers (include/mi_producers.h) for the same
 * system (any pointer may be NULL):
 *   rest_positions [3 block_rows]   the undeformed grid
 *   mass [block_rows]               lumped vertex masses (the kinetic blocks m_i I)
 *   contact_idx [4 n_contacts], contact_g [12 n_contacts], contact_kappa [n_contacts]
 *                                   the contact stencils (vertex, face) and their kappa g g^T data */
int mi_synth_fem_aux(const mi_synth_fem_desc* desc,
                     double*                  rest_positions,
                     double*                  mass,
                     int32_t*                 contact_idx,
                     double*                  contact_g,
                     double*                  contact_kappa);

const char* mi_synth_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* MI_SYNTH_H */
