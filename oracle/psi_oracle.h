/*
 * psi_oracle.h -- TEST INFRASTRUCTURE ONLY (the oracle).
 * FP64 CPU restatement of PSI.voxels(mode='density'); see psi_oracle.c.
 * Only tests/, the smoke check and the CPU-baseline leg of bench.py use it.
 */
#ifndef PSI_ORACLE_H
#define PSI_ORACLE_H
#include <stdint.h>

/* field slots, same tags as the reference (include/grid.h:37-42) */
enum { ORC_M = 0, ORC_X = 1, ORC_V = 2, ORC_XX = 3, ORC_XV = 4, ORC_VV = 5 };

typedef struct {
    int elemtype; /* 4 (tet), 8 (trilinear), 27 (triquadratic) */
    int64_t npart, nelem;
    const double* pos;   /* npart x 3 */
    const double* vel;   /* npart x 3 */
    const double* mass;  /* nelem */
    const int32_t* conn; /* nelem x elemtype */
    int periodic;
    double box[2][3];
} orc_mesh;

typedef struct {
    int n[3];
    double win[2][3];
    double d[3];
    double* f[6]; /* m, x, v, xx, xv, vv; NULL when absent */
} orc_grid;

typedef struct {
    int64_t nfrag;      /* kept (ghost, voxel) leaves, m0 > 0 */
    int64_t ntets;      /* tets after refinement */
    int64_t nghosts;    /* (tet, periodic copy) pairs voxelized */
    int64_t nelem_used; /* elements that passed the window test */
    int maxverts;       /* largest polyhedron produced by a split */
    int maxstack;       /* deepest DFS stack */
    int overflow;       /* vertex or stack overflows (0 on sane input) */
} orc_stats;

int orc_voxels(const orc_mesh* mesh, orc_grid* g, double reftol, int maxlvl, int xvfix, orc_stats* st);
int orc_voxels_range(const orc_mesh* mesh, orc_grid* g, double reftol, int maxlvl, int xvfix, int64_t e0,
                     int64_t e1, orc_stats* st);
int orc_voxels_omp(const orc_mesh* mesh, orc_grid* g, double reftol, int maxlvl, int xvfix, int nthreads, int64_t e0,
                   int64_t e1, orc_stats* st);
int64_t orc_list_fragments(const orc_mesh* mesh, const orc_grid* g, double reftol, int maxlvl, int64_t e0, int64_t e1,
                           int64_t* gi, double* mm, int64_t cap);
int orc_count_tets(const orc_mesh* mesh, const orc_grid* g, double reftol, int maxlvl, int32_t* counts);
int orc_voxelize_one_tet(const orc_mesh* mesh, orc_grid* g, const double* pos12, const double* vel12, double mass,
                         int xvfix, orc_stats* st);

/* point queries: elements whose wrapped box (window test as the R*-tree
 * build) meets the closed box q, ascending; returns the count, writes up to
 * cap indices */
int64_t orc_elem_query(const orc_mesh* m, const double win[2][3], const double q[2][3], int32_t* out, int64_t cap);
/* the multistream sample at x (PSI.VDF); returns the stream count, writes
 * up to cap (rho, vel[3]) */
int64_t orc_sample_vdf(const orc_mesh* m, const double win[2][3], const double* x, double reftol, int maxlvl,
                       double* rho, double* vel, int64_t cap);

/* mode='annihilation': adds the pairwise tet-overlap rate rho0 rho1 vol into
 * the mass field (stats: nfrag = leaves deposited, ntets = tet pairs) */
int orc_annihilation(const orc_mesh* mesh, orc_grid* g, double reftol, int maxlvl, orc_stats* st);

/* mass adjoint (psi_adjoint.c), refine_max_lvl = 0: dpos = dL/dpos
 * (npart x 3, overwritten) from w = dL/dm per cell */
int orc_mass_adjoint(const orc_mesh* mesh, const orc_grid* g, const double* w, double* dpos);
/* adjoint of the fields m, v, vv: dpos, dvel (npart x 3, overwritten, either NULL) from the
 * cells' weights W (13 per cell: dL/dM, dL/dSv[3], dL/dSvv[3][3]); refine level 0; serial */
int orc_adjoint(const orc_mesh* mesh, const orc_grid* g, const double* W, double* dpos, double* dvel);
/* central difference of sum_C G_m m + G_v . v + G_vv : vv for a position (which 0) or velocity
 * (which 1) component of particle p; base: raw sums per cell (M, Sv[3], Svv sym 6) */
int orc_fd_fields_loss(orc_mesh* mesh, const orc_grid* g, const double* base, const double* G, const int64_t* elems,
                       int nel, int64_t p, int a, int which, double h, double* out);
/* central difference (L(x+h) - L(x-h)) / 2h of L = sum (m/V - rho_t)^2 for
 * particle p, axis a; r = m/V - rho_t of the unperturbed mesh; elems are
 * the elements using p */
int orc_fd_mass_loss(orc_mesh* mesh, const orc_grid* g, const double* r, const int64_t* elems, int nel, int64_t p,
                     int a, double h, double* out);

#endif
