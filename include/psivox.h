/*
 * psivox.h -- C ABI of the MI355X voxelizer behind PSI.voxels().
 *
 * Plain C types and pointers only.  The Python module PSI (pypsi_amd/csrc/
 * psimodule.c) calls these; so can any other host language through its FFI.
 *
 * Reference interface replaced:
 *   void psi_voxels(psi_grid* grid, psi_mesh* mesh, psi_rtree* rtree,
 *                   psi_int mode, psi_real reftol, psi_int max_ref_lvl);
 *     -- src/cpsi.c:41 (forward declared src/PSI.c:39, called src/PSI.c:1015
 *        with rtree = NULL), density branch src/cpsi.c:100-112.
 *   psi_grid (include/grid.h:46-53) -> psivox_grid
 *   psi_mesh (include/mesh.h:42-52) -> psivox_mesh
 *
 * All entry points return 0 on success and a negative PSIVOX_E* code on
 * failure; psivox_last_error() then gives a message (thread-local).
 * Results are deposited into the grid's existing contents (the reference
 * never clears the grid: src/cpsi.c:41-175, src/geometry.c:164-186).
 */
#ifndef PSIVOX_H
#define PSIVOX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PSIVOX_ABI_VERSION 2

/* element types: vertices per element (include/mesh.h:26-28) */
#define PSIVOX_SIMPLEX 4
#define PSIVOX_LINEAR 8
#define PSIVOX_QUADRATIC 27

/* field slots, the reference's tags (include/grid.h:37-42) */
#define PSIVOX_M 0
#define PSIVOX_X 1
#define PSIVOX_V 2
#define PSIVOX_XX 3
#define PSIVOX_XV 4
#define PSIVOX_VV 5

/* modes (include/psi.h:71-72).  PSIVOX_MODE_ANNIHILATION adds the integral
 * of rho^2 over each cell into the mass field (src/cpsi.c:115-158): every
 * ordered pair of tets whose boxes overlap adds rho0 rho1 vol(overlap ∩
 * cell), FP64, whatever the kernel option; the other fields are untouched. */
#define PSIVOX_MODE_DENSITY 0
#define PSIVOX_MODE_ANNIHILATION 1

/* kernels */
#define PSIVOX_KERNEL_AUTO 0  /* the column kernel: FP32 for most tets, FP64 for tets spanning more than
                                 4 cells along an axis (their smallest fragments need it) */
#define PSIVOX_KERNEL_PLAIN 1 /* one thread per tet, FP64 clip */
#define PSIVOX_KERNEL_FAST 2  /* one lane per (tet, column) task, FP32 cell-local clip for every tet, FP64 deposit */
#define PSIVOX_KERNEL_FAST64 3 /* the column kernel in FP64 for every tet */

/* error codes */
#define PSIVOX_OK 0
#define PSIVOX_E_ARG -1      /* invalid argument (message says which) */
#define PSIVOX_E_NODEV -2    /* no usable GPU */
#define PSIVOX_E_HIP -3      /* a HIP runtime call failed */
#define PSIVOX_E_NOMEM -4    /* device allocation failed */
#define PSIVOX_E_OVERFLOW -5 /* a polyhedron exceeded its vertex budget */
#define PSIVOX_E_MODE -6     /* mode not supported by this build */
#define PSIVOX_E_CANCELLED -7 /* the progress callback asked to stop */

/* Mesh: same arrays as psi_mesh (include/mesh.h:42-52).  pos/vel are
 * npart x 3 float64, mass has one entry per element, conn is
 * nelem x elemtype int32 in the reference's corner order (4i+2j+k for
 * trilinear elements, 9i+3j+k for triquadratic ones). vel may be NULL
 * (zero velocities). */
typedef struct {
    int32_t elemtype;
    int32_t periodic; /* 1: positions wrap in box (src/PSI.c:383-396) */
    int64_t npart;
    int64_t nelem;
    const double* pos;
    const double* vel;
    const double* mass;
    const int32_t* conn;
    double box[2][3];
    int32_t on_device; /* 1: the four pointers are device pointers */
    int32_t reserved;
} psivox_mesh;

/* Grid: a Cartesian window split into n cells, psi_grid (include/grid.h:
 * 46-53).  fields[k] point to C-ordered float64 arrays in the reference's
 * layout (src/PSI.c:119-186): m (n0,n1,n2), x and v (n0,n1,n2,3), xx, xv, vv
 * (n0,n1,n2,3,3); NULL for fields not requested.  m is required. */
typedef struct {
    int32_t n[3];
    int32_t on_device; /* 1: fields are device pointers */
    double window[2][3];
    double d[3]; /* cell size; (window[1]-window[0])/n as src/PSI.c:111-113 */
    double* fields[6];
} psivox_grid;

/* Called between element batches with the number of elements done; a nonzero
 * return cancels the call (PSIVOX_E_CANCELLED), as the reference's per-element
 * PyErr_CheckSignals() does (src/cpsi.c:160-163).  Fields then hold the
 * batches deposited so far. */
typedef int (*psivox_progress_fn)(void* ctx, int64_t elements_done, int64_t elements_total);

typedef struct {
    int32_t mode;             /* PSIVOX_MODE_DENSITY */
    int32_t refine_max_lvl;   /* clamped to 6 (src/refine.c:89-92) */
    double refine_tolerance;  /* in grid cells (src/cpsi.c:63-66) */
    int32_t kernel;           /* PSIVOX_KERNEL_* */
    int32_t xv_true_covariance; /* 0: reference xv formula (src/geometry.c:139-140) */
    void* stream;             /* hipStream_t to run on; NULL = library stream */
    psivox_progress_fn progress; /* may be NULL */
    void* progress_ctx;
    int64_t batch_elements;   /* elements per device batch; 0 = automatic */
} psivox_opts;

typedef struct {
    int64_t nelem_used; /* elements inside the window after periodic wrap */
    int64_t ntets;      /* tets after refinement */
    int64_t nghosts;    /* (tet, periodic copy) pairs; annihilation: tet pairs voxelized */
    int64_t nfrag;      /* (copy, voxel) fragments deposited, m0 > 0 */
    double kernel_ms;   /* device time of the voxelization kernel: the FP64 kernel, or the
                           column kernel (task setup, ordering and residual kernels not
                           included); a batched call reports its whole column pass */
    double total_ms;    /* wall time of the whole call */
    int32_t kernel;     /* kernel that ran (PSIVOX_KERNEL_FAST for the column kernel in either precision) */
    int32_t reserved;   /* 1: a tet the column path cannot take sent the call to the FP64 kernel */
    int64_t nwide;      /* tets the automatic choice gave to the FP64 column kernel */
} psivox_stats;

/* One-mesh voxelization, PSI.voxels(grid, mesh, mode, refine_tolerance,
 * refine_max_lvl) -- src/PSI.c:974-1018 -> src/cpsi.c:41-175. */
int psivox_voxels(const psivox_mesh* mesh, psivox_grid* grid, const psivox_opts* opts, psivox_stats* stats);

/* Many independent (mesh, grid) pairs in one device pass.  No reference
 * counterpart: equivalent to calling psi_voxels (src/cpsi.c:41) once per
 * pair.  stats may be NULL or point to nmesh entries. */
int psivox_voxels_batched(int32_t nmesh, const psivox_mesh* meshes, psivox_grid* grids, const psivox_opts* opts,
                          psivox_stats* stats);

/* psivox_voxels without allocating device memory.  No counterpart in the
 * reference: it is for callers that cannot allocate (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The mesh
 * and the grid must be device arrays (on_device = 1), mode density,
 * refine_max_lvl 0.  Every buffer beyond them comes from a fixed 256 MiB
 * scratch array in device memory that each GPU holds until the process
 * ends; the grid's raw-moment accumulator (nvox x 1/10/28 doubles for m,
 * m+v+vv, all fields) must leave room in it for one element batch and one
 * chunk of column tasks, else PSIVOX_E_NOMEM with a message.  Calls on one
 * GPU share the array, so calls that overlap on one GPU race; the library
 * serializes the calls of one process.  Runs on opts->stream (NULL: the
 * null stream); stats->kernel_ms is 0 (no events are recorded). */
int psivox_voxels_noalloc(const psivox_mesh* mesh, psivox_grid* grid, const psivox_opts* opts, psivox_stats* stats);

/* Adjoint of the mass deposit.  No reference counterpart (the reference has
 * no derivatives; the forward it differentiates is psi_voxels,
 * src/cpsi.c:41, density branch :100-112).  Given dL/dm per cell (float64,
 * shaped like grid m), ADDS dL/dpos (npart x 3 float64) of the mass field m
 * that psivox_voxels produces on the same mesh and grid (with the same
 * refine_tolerance and refine_max_lvl) into dl_dpos.  On refined meshes a
 * tet's vertices are the element's trilinear or triquadratic map at fixed
 * parameter points, so their gradients reach the element's nodes through
 * that map; the refinement decisions are piecewise constant and carry no
 * gradient.  dl_dm is a device pointer when grid->on_device, dl_dpos when
 * mesh->on_device; grid->fields is not used. */
int psivox_mass_adjoint(const psivox_mesh* mesh, const psivox_grid* grid, const double* dl_dm, double* dl_dpos,
                        const psivox_opts* opts);

/* Adjoint of the fields m, v and vv.  No reference counterpart (the
 * forward it differentiates is psi_voxels, src/cpsi.c:41, density branch
 * :100-112, with the deposit of src/geometry.c:106-186).  grid->fields
 * holds the forward's m, v and vv (the fields psivox_voxels wrote into
 * zeroed arrays; v and vv may be NULL when their gradients are).
 * dl_dfields[6] points to dL/d(m, x, v, xx, xv, vv) in the fields' layouts
 * (NULL: zero; x, xx and xv must be NULL).  ADDS dL/dpos and dL/dvel
 * (npart x 3 float64 each; either may be NULL), on refined meshes as
 * psivox_mass_adjoint.  Device pointers follow grid->on_device (fields and
 * their gradients) and mesh->on_device (dl_dpos, dl_dvel).  Position
 * gradients: the plane-section integrals of the mass adjoint with the
 * velocity fields' cubic integrand; velocity gradients: the fragments'
 * q-moments against each cell's weights.  Errors: PSIVOX_E_MODE for
 * gradients of x, xx or xv. */
int psivox_adjoint(const psivox_mesh* mesh, const psivox_grid* grid, const double* const* dl_dfields, double* dl_dpos,
                   double* dl_dvel, const psivox_opts* opts);

/* ---- point queries: RStarTree and PSI.VDF ---- */

/* An index of the mesh's element boxes on the GPU.  Replaces the R*-tree
 * (psi_rtree, include/rtree.h:40-71) that psi_rtree_from_mesh builds
 * (src/rtree.c:48-73): every element's vertex box after the periodic wrap of
 * psi_aabb_periodic (src/cpsi.c:309-341), elements that test outside the
 * window left out.  The index keeps a device copy of a host mesh (a
 * snapshot: later changes to the host arrays are not seen); a device mesh's
 * arrays must outlive the index.  Build, query and sample on the device that
 * was current at build time; stream may be NULL (library stream). */
typedef struct psivox_elem_index psivox_elem_index;

int psivox_elem_index_build(const psivox_mesh* mesh, const double window[2][3], void* stream,
                            psivox_elem_index** out);
void psivox_elem_index_free(psivox_elem_index* idx);

/* The elements whose box meets the closed box q -- RStarTree.query
 * (src/PSI.c:817-861; psi_rtree_query_init/next, src/rtree.c:75-123) -- in
 * ascending element order (the reference returns them in its tree's order).
 * Writes min(*n, cap) indices to the host array out and the count to *n. */
int psivox_elem_index_query(psivox_elem_index* idx, const double q[2][3], int32_t* out, int64_t cap, int64_t* n,
                            void* stream);

/* The multistream sample at a point -- psi_sample_vdf (src/cpsi.c:177-275),
 * PSI.VDF (src/PSI.c:1022-1058).  Each element whose box holds the point is
 * refined on its raw vertex positions as psi_voxels refines (tolerance in
 * position units, src/cpsi.c:197; level clamped to 6); every tet holding the
 * point (barycentric coordinates >= 0, nonzero volume; src/geometry.c:335-389)
 * is one stream: rho = tet mass / tet volume, vel = barycentric interpolation
 * of its vertex velocities.  Streams in ascending element order, tets in
 * refinement order.  Writes min(*n, cap) streams (rho[i]; vel[3i..3i+3)) to
 * host arrays and the count to *n. */
int psivox_sample_vdf(psivox_elem_index* idx, const double point[3], double refine_tolerance,
                      int32_t refine_max_lvl, double* rho, double* vel, int64_t cap, int64_t* n, void* stream);

const char* psivox_last_error(void);
int psivox_abi_version(void);
/* number of visible GPUs (0 if none); never throws */
int psivox_device_count(void);

#ifdef __cplusplus
}
#endif
#endif
