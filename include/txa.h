/*
 * txa.h -- C ABI of the MI355X finite-element assembly engine.
 *
 * The engine replaces, for one GPU, Tianxin/Panzer's volume residual +
 * Jacobian path: workset geometry (IntegrationValues2, BasisValues2), gather,
 * DOF/DOFGradient, the integrator evaluators, the Sacado-FAD element Jacobian
 * (here an explicit tangent) and ScatterResidual_Tpetra into a Tpetra-layout
 * CRS matrix addressed through GlobalIndexer local IDs.
 *
 * Conventions
 *  - Every entry point returns TXA_OK (0) or an error code; nothing throws
 *    across the ABI.  txa_last_error() gives the message of the last failure on
 *    a context.  (The reference throws std::logic_error / std::invalid_argument
 *    via TEUCHOS_TEST_FOR_EXCEPTION; the C++ shim in txa_panzer.hpp rethrows.)
 *  - Borrow-only: pointers passed in are read during the call (setup calls) or
 *    until the stream reaches the launched work (assembly calls).  Mesh,
 *    DOF-map and graph arrays are owned by the context and live on the device
 *    until txa_mesh_destroy / txa_destroy.
 *  - "host" pointers are CPU memory, "device" pointers are HIP device memory of
 *    the context's device.  `stream` is a hipStream_t (NULL = default stream).
 *  - A context is not thread-safe; calls on it are stream-ordered.
 */
#ifndef TXA_H
#define TXA_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define TXA_OK 0
#define TXA_ERR_INVALID 1     /* bad argument (std::invalid_argument in the reference) */
#define TXA_ERR_LOGIC 2       /* inconsistent state (std::logic_error in the reference) */
#define TXA_ERR_DEVICE 3      /* HIP runtime failure */
#define TXA_ERR_NOMEM 4       /* device allocation failed */
#define TXA_ERR_UNSUPPORTED 5 /* combination not implemented */

/* Side sets of the brick: the shards hex side ordinals the cube mesh uses
 * (adapters-stk/src/stk_interface/Panzer_STK_CubeHexMeshFactory.cpp:660-760). */
#define TXA_SIDE_BOTTOM 0 /* y = 0 */
#define TXA_SIDE_RIGHT 1  /* x = max */
#define TXA_SIDE_TOP 2    /* y = max */
#define TXA_SIDE_LEFT 3   /* x = 0 */
#define TXA_SIDE_BACK 4   /* z = 0 */
#define TXA_SIDE_FRONT 5  /* z = max */

/* element topologies (shards: Hexahedron<8>, Tetrahedron<4> from a 6-tet split
 * of each hex, HGRAD_HEX_C2 on Hexahedron<8> geometry) */
#define TXA_HEX8 1
#define TXA_TET4 2
#define TXA_HEX27 3
#define TXA_WEDGE6 4   /* shards Wedge<6>: triangle 0,1,2, then 3,4,5 above it (HGRAD_WEDGE_C1) */

/* physics */
#define TXA_DIFFUSION 0       /* k grad(u).grad(v) + rho_c udot v - f v   (scalar) */
#define TXA_ELASTICITY 1      /* sigma(u):eps(v) + rho (a + a_R v).v + b_R sigma(v):eps(v) - b.v */

/* element-integration precision (geometry and scatter-accumulation are FP64 in all) */
#define TXA_FP64 0
#define TXA_FP32 1
#define TXA_SPLIT_BF16 2

typedef struct txa_context *txa_ctx;

/* Context on one HIP device.
 * Replaces the per-rank Kokkos execution space + Tpetra node that
 * panzer::AssemblyEngine runs on (disc-fe/src/Panzer_AssemblyEngine_impl.hpp:67-129). */
int txa_create(int device, txa_ctx *out);
int txa_destroy(txa_ctx ctx);
const char *txa_last_error(txa_ctx ctx);
/* Waits for the context's device work.  Returns TXA_ERR_INVALID when an assembly or
 * batch since the previous txa_sync was given a Dirichlet LID outside its DOF map, or a
 * txa_neumann_sides call a side that does not exist (those entries touched nothing; the
 * reference throws for them), and clears the report. */
int txa_sync(txa_ctx ctx);
/* Library version string (static). */
const char *txa_version(void);

/* Structured brick mesh: nx*ny*nz hex cells (x fastest), vertex id
 * i + (nx+1)(j + (ny+1)k), element order and local vertex order of
 * CubeHexMeshFactory::buildBlock
 * (adapters-stk/src/stk_interface/Panzer_STK_CubeHexMeshFactory.cpp:401-460) and,
 * for TXA_TET4 / TXA_HEX27, the entity numbering of CartesianConnManager
 * (dof-mgr/test/cartesian_topology/CartesianConnManager.cpp:220-235, 744-890).
 * vertex_coords: host [(nx+1)(ny+1)(nz+1)][3] FP64 (any geometry with that
 * topology).  Replaces STK mesh construction + STKConnManager::buildConnectivity
 * (adapters-stk/src/Panzer_STKConnManager.cpp:243-300). */
int txa_mesh_brick(txa_ctx ctx, int topo, int nx, int ny, int nz, const double *vertex_coords,
                   int *mesh_id);
/* A general mesh (e.g. read from Exodus, or a ConnManager's connectivity):
 * conn host [nelem][npe] entity ids, npe = 8 (TXA_HEX8, shards vertex order),
 * 4 (TXA_TET4), 27 (TXA_HEX27: 8 vertices, then 12 edges, 6 faces and the cell,
 * as in the brick).  Vertex entries are vertex ids in [0, nverts); the others
 * any ids below nentities (nentities = nverts for vertex-only topologies).
 * vertex_coords host [nverts][3].  The DOF map and graph follow the same rules as
 * for a brick.  The assembly of a single-block mesh takes the node-patch kernels
 * (one-field hex8 / tet4 diffusion, FP64 or FP32; see txa_mesh_patches) or the
 * element kernels with a general node gather (3-field hex8 / hex27 elasticity, every
 * precision); other meshes take the generic kernels in FP64.  Brick side sets
 * (TXA_SIDE_*) and batches need a brick; listed sides (txa_neumann_sides) do not.
 * Replaces the connectivity STKConnManager / ConnManager hands to the DOF manager
 * (adapters-stk/src/Panzer_STKConnManager.cpp:243-300). */
int txa_mesh_upload(txa_ctx ctx, int topo, int64_t nelem, const int64_t *conn, int64_t nentities, int64_t nverts,
                    const double *vertex_coords, int *mesh_id);
/* One element block of a mesh: elements of one topology (TXA_HEX8, TXA_TET4, TXA_WEDGE6 or
 * TXA_HEX27) with their connectivity, host [nelem][npe] as for txa_mesh_upload. */
typedef struct {
  int topo;
  int64_t nelem;
  const int64_t *conn;
} txa_element_block;

/* A mesh of several element blocks, of one or of different topologies, sharing the
 * vertices (and entity ids) where they meet: one DOF map and one CRS graph over all of
 * them.  Elements are numbered block after block in the order given (block b's element
 * e is element offset_b + e: coefficient arrays run over that numbering), and the first
 * touch that orders the LIDs walks the blocks in that order, as DOFManager walks its
 * blockOrder_ (dof-mgr/src/Panzer_DOFManager.cpp:604-628).  txa_dofmap_copy's element
 * LIDs are the blocks' [nelem][npe * nfields] arrays one after another.  Assembly runs
 * each block with its own basis and cubature and sums into the same rows, as
 * AssemblyEngine evaluates each block's volume field manager into one container
 * (disc-fe/src/Panzer_AssemblyEngine_impl.hpp:134-180; element blocks of different
 * topology: adapters-stk/src/Panzer_STK_WorksetFactory.cpp:279-292). */
int txa_mesh_upload_blocks(txa_ctx ctx, int nblocks, const txa_element_block *blocks, int64_t nentities,
                           int64_t nverts, const double *vertex_coords, int *mesh_id);
/* out (host, 1 + 3 * nblocks entries): out[0] = blocks, then per block its topology,
 * element count and first element.  cap: entries available in out. */
int txa_mesh_blocks(txa_ctx ctx, int mesh_id, int64_t *out, int64_t cap);
int txa_mesh_destroy(txa_ctx ctx, int mesh_id);
/* out[0]=elements out[1]=connectivity entries/element (0 when the blocks differ) out[2]=vertices */
int txa_mesh_info(txa_ctx ctx, int mesh_id, int64_t *out);
/* *affine = 1 when the mesh is a HEX8 brick whose cells are all parallelepipeds (the
 * bilinear and trilinear terms of every cell's map below 2^-47 of its coordinate
 * magnitude, i.e. at the rounding of the coordinates), decided once when the brick is
 * created: its lattice kernels then form one metric per cell instead of one per
 * integration point.  0 otherwise.  Introspection only; no counterpart in the reference. */
int txa_mesh_affine(txa_ctx ctx, int mesh_id, int *affine);
/* Node patches of an uploaded one-field HEX8 / TET4 mesh of one block, built with its graph:
 * out[0] = rows per patch, out[1] = patches, out[2] = element instances (elements counted
 * once per patch they touch), out[3] = element colours, out[4] = longest row, out[5] = 1 when
 * the state is read in curve order (the LIDs do not follow the curve; a gather pass precedes
 * each assembly), 0 when it is read by LID, out[6] = 1 when every hex8 element is a
 * parallelepiped (the one-metric element form).  out holds 7 entries.  All zero
 * when the mesh has none (bricks, several blocks, other topologies or fields, or more than
 * 64 colours): its assembly then takes the element kernels.  Introspection only; no
 * counterpart in the reference. */
int txa_mesh_patches(txa_ctx ctx, int mesh_id, int64_t *out);

/* DOF map for `nfields` fields on every connectivity entity (1 = scalar,
 * 3 = displacement).  GIDs by the GUN rule on one rank, owned_ by first touch
 * over elements, LIDs = position in owned_.  Built on the device.
 * Replaces DOFManager::buildGlobalUnknowns (dof-mgr/src/Panzer_DOFManager.cpp:449-715)
 * and GlobalIndexer::buildLocalIds (dof-mgr/src/Panzer_GlobalIndexer.hpp:548-640). */
int txa_dofmap_build(txa_ctx ctx, int mesh_id, int nfields, int64_t *ndofs);
/* Copies (host, nullable): owned GIDs [ndofs]; element LIDs [elements][ndof_per_elem]
 * (GlobalIndexer::getOwnedIndices / getElementLIDs, Panzer_GlobalIndexer.hpp:254-320). */
int txa_dofmap_copy(txa_ctx ctx, int mesh_id, int64_t *owned_gids, int32_t *elem_lids);
/* LID of every DOF-carrying lattice point and field, host [points][nfields]:
 * vertex i + (nx+1)(j + (ny+1)k) for HEX8/TET4, Q2 node I + (2nx+1)(J + (2ny+1)K)
 * for HEX27.  The nodal form of getElementLIDs + getGIDFieldOffsets
 * (dof-mgr/src/Panzer_GlobalIndexer.hpp:254-320, dof-mgr/src/Panzer_DOFManager.cpp:376-389),
 * used to address nodesets (Dirichlet rows, TianXin_PointEvaluator_impl.hpp:44-190). */
int txa_dofmap_point_lids(txa_ctx ctx, int mesh_id, int32_t *point_lids, int64_t *npoints);

/* CRS graph in LID space: rows = LIDs, columns = union of element closures,
 * sorted ascending, duplicates merged.  Built on the device.
 * Replaces TpetraLinearObjFactory::buildGhostedGraph + fillComplete
 * (disc-fe/src/lof/Panzer_TpetraLinearObjFactory_impl.hpp:559-650). */
int txa_graph_build(txa_ctx ctx, int mesh_id, int64_t *nnz);
int txa_graph_copy(txa_ctx ctx, int mesh_id, int64_t *rowptr, int32_t *col);   /* host outputs */
int txa_graph_device(txa_ctx ctx, int mesh_id, const int64_t **rowptr, const int32_t **col);

/* The assembled system in the MatrixMarket exchange format: A (CRS values vals
 * on the mesh's graph) as "coordinate real general", f and x as "array real
 * general".  vals, f and x are device arrays; a null array or path skips that
 * file.  Indices are GID + 1, rows in LID order, values in the shortest
 * decimal form that reads back to the same double.
 * Replaces TpetraLinearObjContainer::writeMatrixMarket
 * (disc-fe/src/lof/Panzer_TpetraLinearObjContainer.hpp:359-367). */
int txa_write_matrix_market(txa_ctx ctx, int mesh_id, const double *vals, const double *f, const double *x,
                            const char *a_path, const char *b_path, const char *x_path);

/* Reads a MatrixMarket "array real general" vector of n entries into the device
 * array dst (entry i to LID i).  TXA_ERR_INVALID for a missing or malformed
 * file, or one of another length.
 * Replaces TpetraLinearObjFactory::readVector
 * (disc-fe/src/lof/Panzer_TpetraLinearObjFactory_impl.hpp:743-772). */
int txa_read_vector_mm(txa_ctx ctx, const char *path, double *dst, int64_t n);

/* One residual + Jacobian evaluation.  All arrays are device memory.
 * f (ndofs) and vals (nnz) are overwritten with the assembled sums, Dirichlet
 * rows applied last (f_i = x_i - g_i, identity matrix row).
 * Jacobian = beta dR/dx + alpha dR/dxdot + gamma2 dR/dxddot, the seeds of
 * GatherSolution_Tpetra<Jacobian>
 * (disc-fe/src/evaluators/Panzer_GatherSolution_Tpetra_impl.hpp:549-612).
 * Replaces AssemblyEngine<Jacobian>::evaluate volume + Dirichlet fill
 * (disc-fe/src/Panzer_AssemblyEngine_impl.hpp:67-177, 305-321) running
 * Gather -> DOF/DOFGradient -> Integrator_* -> ScatterResidual_Tpetra
 * (disc-fe/src/evaluators/Panzer_ScatterResidual_Tpetra_impl.hpp:375-520). */
typedef struct {
  int physics;                 /* TXA_DIFFUSION or TXA_ELASTICITY */
  int precision;               /* TXA_FP64, TXA_FP32, TXA_SPLIT_BF16 */
  const double *coef;          /* [elements] k (diffusion) or Young's modulus (elasticity) */
  double nu;                   /* Poisson ratio */
  double rho;                  /* rho*c (diffusion) or density (elasticity) */
  double src[3];               /* source f (diffusion) / body force per volume (elasticity) */
  double rayleigh_a, rayleigh_b;
  const double *x, *xdot, *xddot;   /* [ndofs] by LID; xdot/xddot may be NULL */
  double alpha, beta, gamma2;
  int transient;               /* workset.evaluate_transient_terms */
  double *f;                   /* [ndofs] residual, or NULL */
  double *vals;                /* [nnz] CRS values, or NULL */
  const int32_t *bc_lids;      /* [n_bc] Dirichlet LIDs, or NULL */
  const double *bc_vals;       /* [n_bc] prescribed values, or NULL (= 0) */
  int64_t n_bc;
  /* Neumann flux on sides of the brick (hex8 and hex27 meshes), added to f before
   * the Dirichlet rows: f(b, i) += neumann_values[k nf + i] * (integral over side
   * neumann_sides[k] of N_b dS), Gauss rule of the volume degree on each face.
   * Sides (TXA_SIDE_*) are the shards hex side ordinals of the cube mesh's side
   * sets.  Host arrays, read during the call.
   * Replaces TianXin::Flux::evaluateFields
   * (disc-fe/src/evaluators/TianXin_Neumann_impl.hpp:141-160) on a side workset. */
  int n_neumann;
  const int *neumann_sides;     /* [n_neumann] */
  const double *neumann_values; /* [n_neumann][nf] flux per field */
  /* Element block to assemble: 0 = every block of the mesh, b > 0 = block b - 1 only (one
   * evaluator per element block, as each block has its own volume field manager). */
  int block;
  /* 0: f and vals are overwritten with this call's sums; 1: the sums are added to what f and
   * vals hold (several blocks or evaluators scattering into one container, which the caller
   * zeroed, as ScatterResidual_Tpetra adds into a container AssemblyEngine initialised). */
  int accumulate;
} txa_assembly_args;

int txa_assemble(txa_ctx ctx, int mesh_id, const txa_assembly_args *args, void *stream);

/* Dirichlet rows alone: f_i = x_i - g_i (g = values or 0) and row i of the CRS values
 * (vals) becomes the identity row, for the n LIDs in lids (device arrays; f or vals may be
 * NULL).  The BC step of AssemblyEngine after every block's volume assembly
 * (disc-fe/src/Panzer_AssemblyEngine_impl.hpp:96-126; TianXin_Dirichlet_impl.hpp:50-82,
 * Panzer_TpetraLinearObjContainer.hpp:228-237).  LIDs outside the map are reported by txa_sync. */
int txa_apply_dirichlet(txa_ctx ctx, int mesh_id, const int32_t *lids, const double *values, int64_t n,
                        const double *x, double *f, double *vals, void *stream);

/* Neumann flux on listed element sides of any mesh (uploaded, multi-block or brick): for side
 * k, the element elems[k] (mesh element index, blocks in order) and its shards side ordinal
 * ords[k] (Hexahedron<8>/<27>: 0..5, Tetrahedron<4>: 0..3, Wedge<6>: 0..2 quads, 3..4
 * triangles), f(node b of the side, field i) += values[k nf + i] * integral over the side of
 * N_b dS: 2x2 Gauss over the bilinear face map for hex8 and wedge quads, 3x3 with the
 * biquadratic face basis for hex27, one third of the area per vertex on triangles.  Device
 * arrays; f gains the loads with FP64 atomics (not bitwise reproducible when sides share
 * nodes).  A side that does not exist touches nothing and is reported by txa_sync.
 * Replaces TianXin::Flux::evaluateFields on the side worksets of a side set
 * (disc-fe/src/evaluators/TianXin_Neumann_impl.hpp:141-160; side worksets of
 * adapters-stk/src/Panzer_STK_WorksetFactory.cpp:150-245: cells with their local side index). */
int txa_neumann_sides(txa_ctx ctx, int mesh_id, int64_t n, const int64_t *elems, const int *ords,
                      const double *values, double *f, void *stream);

/* p[i] = value for the n doubles of device array p: zeroing a container's f and matrix
 * values before an assembly accumulates into them (AssemblyEngine's initialisation of the
 * ghosted container, Thyra::assign / initializeMatrix in
 * disc-fe/src/Panzer_ModelEvaluator_impl.hpp:1785-1790). */
int txa_fill(txa_ctx ctx, double *p, int64_t n, double value, void *stream);

/* y = A x for the CRS values vals on the mesh's graph (device arrays of ndofs / nnz
 * entries): Tpetra::CrsMatrix::apply, the operator a linear solver applies to W. */
int txa_spmv(txa_ctx ctx, int mesh_id, const double *vals, const double *x, double *y, void *stream);

/* ModelEvaluator-shaped evaluation: the hand-off to a nonlinear solver.  A model binds a
 * mesh (with its DOF map and graph), the physics and material of txa_assembly_args and the
 * Dirichlet rows; txa_model_eval then fills the residual f and the Jacobian W as
 * panzer::ModelEvaluator::evalModelImpl_basic does
 * (disc-fe/src/Panzer_ModelEvaluator_impl.hpp:725-780, 1715-1840):
 *  - a model without transient support always evaluates the steady system (alpha 0, beta 1);
 *  - with transient support alpha, beta (and gamma2 for the second derivative) come from
 *    the in-arguments, and the transient terms are evaluated when x_dot is given;
 *  - f and W are outputs each may leave NULL: f and W, f alone (the Residual evaluation),
 *    or W alone (the Jacobian evaluation); both NULL does nothing.
 * W = alpha df/dx_dot + beta df/dx + gamma2 df/dx_dot_dot, the CRS values on the mesh's graph.
 * The Dirichlet rows are applied last (f_i = x_i - g_i, identity rows). */
typedef struct {
  int physics, precision;      /* as txa_assembly_args */
  const double *coef;          /* device [elements] */
  double nu, rho;
  double src[3];
  double rayleigh_a, rayleigh_b;
  int transient_support;       /* build_transient_support of the reference's ModelEvaluator */
  const int32_t *bc_lids;      /* device [n_bc] Dirichlet LIDs, or NULL */
  const double *bc_vals;       /* device [n_bc] prescribed values, or NULL (= 0) */
  int64_t n_bc;
} txa_model_desc;
typedef struct {
  const double *x;             /* device [ndofs] */
  const double *x_dot;         /* device [ndofs] or NULL (steady fill) */
  const double *x_dot_dot;     /* device [ndofs] or NULL */
  double alpha, beta, gamma2;
  double t, step_size;         /* time and step of the evaluation (coefficients are time independent here) */
} txa_model_in;
typedef struct {
  double *f;                   /* device [ndofs] or NULL */
  double *W;                   /* device [nnz] CRS values or NULL */
} txa_model_out;
int txa_model_create(txa_ctx ctx, int mesh_id, const txa_model_desc *desc, int *model_id);
int txa_model_eval(txa_ctx ctx, int model_id, const txa_model_in *in, const txa_model_out *out, void *stream);
int txa_model_destroy(txa_ctx ctx, int model_id);

/* txa_assemble without any allocation.  No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  The assembly workspace (the
 * element blocks of the two-phase elasticity kernels, the element matrices of the
 * multi-block row gather, the curve-ordered state of a scattered node-patch mesh)
 * comes from one static device array of 128 MiB per
 * device instead of the context's pool; an assembly that needs more is refused with
 * TXA_ERR_NOMEM before anything is launched.  Results equal txa_assemble's.  Every
 * call on one device shares that array: calls whose work overlaps on one GPU (other
 * streams, other contexts, other threads) race and must be ordered by the caller. */
int txa_assemble_noalloc(txa_ctx ctx, int mesh_id, const txa_assembly_args *args, void *stream);

/* Batched assembly of many resident meshes (HEX8 and TET4, Q1 diffusion) in one
 * persistent launch: every mesh contributes its tiles to one work list that the
 * launch drains, then one Dirichlet launch covers all meshes.  With use_graph
 * the two launches (and the upload of the per-mesh argument blocks) are
 * captured once into a HIP graph and replayed.  Replaces the element-block x
 * workset double loop of AssemblyEngine::evaluateVolume
 * (disc-fe/src/Panzer_AssemblyEngine_impl.hpp:134-180): one batch = every block
 * of every model, one launch instead of one evaluateFields per workset.
 * Meshes of a batch need a scalar DOF map and a graph and cannot be destroyed
 * while the batch exists.  All items share `precision` (TXA_FP64 or TXA_FP32)
 * and the transient flag. */
int txa_batch_create(txa_ctx ctx, const int *mesh_ids, int n, int precision, int transient, int *batch_id);
int txa_batch_destroy(txa_ctx ctx, int batch_id);
/* args: host array [n], one per mesh in creation order (device pointers inside,
 * as for txa_assemble).  Host-blocking only while a previous launch of the
 * same batch still reads its argument blocks. */
int txa_batch_assemble(txa_ctx ctx, int batch_id, const txa_assembly_args *args, int use_graph, void *stream);
/* Autotuner: times `reps` launches for each tile shape (32 x 8 or 16 x 16 vertex
 * columns per workgroup), slab height (vertex layers per work item: 8, 16, 32,
 * 64, whole column) and persistent grid (1, 2, 4 workgroups per CU), keeps the
 * fastest and reports its time per launch in
 * best_ms.  The counterpart of the workset-size sweep in
 * adapters-stk/example/MixedPoissonExample/analyze_performance.py:57-61; it
 * writes f / vals as an assembly does. */
int txa_batch_tune(txa_ctx ctx, int batch_id, const txa_assembly_args *args, int reps, void *stream,
                   double *best_ms);
/* Sets the work shape by hand instead of txa_batch_tune: tile width (16 or 32 vertex
 * columns along x; the tile is tile_x x 256/tile_x), slab height (vertex layers per work
 * item, 0 = whole column) and persistent workgroups.  The counterpart of the "Workset
 * Size" entry of the assembly parameters
 * (adapters-stk/src/Panzer_STK_ModelEvaluatorFactory_impl.hpp:294). */
int txa_batch_shape(txa_ctx ctx, int batch_id, int tile_x, int slab, int grid);
/* out[0]=meshes out[1]=work items out[2]=slab height (0 = whole) out[3]=persistent workgroups
 * out[4]=tile width (vertex columns along x; the tile is out[4] x 256/out[4])
 * out[5]=kernel launches per call besides the Dirichlet pass (1: both element kinds in one) */
int txa_batch_info(txa_ctx ctx, int batch_id, int64_t *out);

/* Serving: a worker thread inside the library admits assembly requests for
 * resident meshes by HBM footprint and runs admitted requests concurrently,
 * one HIP stream per slot; requests that arrive while others run are admitted
 * as slots and memory free up (continuous batching).  Each request is one
 * implicit time step's residual + Jacobian, the evaluate() call of
 * AssemblyEngine with the time-integration coefficients of
 * AssemblyEngineInArgs (disc-fe/src/Panzer_AssemblyEngine_impl.hpp:67-129,
 * Panzer_AssemblyEngine_InArgs.hpp:95-106).  Second-order problems use
 * Newmark: R = M a + C v + K u - f with C = rayleigh_a M + rayleigh_b K and
 * J = K + newmark_gamma/(newmark_beta dt) C + 1/(newmark_beta dt^2) M;
 * first-order problems use J = K + (1/dt) M. */
typedef struct {
  int mesh_id;                 /* a mesh the server was created with */
  int physics;                 /* TXA_ELASTICITY (Newmark) or TXA_DIFFUSION */
  int precision;               /* as txa_assembly_args */
  const double *coef;          /* device [elements] */
  double nu, rho;
  double src[3];
  double rayleigh_a, rayleigh_b;
  double dt, newmark_beta, newmark_gamma;
  const double *u, *v, *a;     /* device [ndofs] state, read until the request is done; a may be NULL for diffusion */
  const int32_t *bc_lids;      /* device [n_bc] or NULL */
  const double *bc_vals;       /* device [n_bc] or NULL (= 0) */
  int64_t n_bc;
  int want_jacobian;           /* 0: residual only */
  void *stream;                /* hipStream_t that produced u/v/a/coef/bc (NULL = the default
                                  stream): the request runs after the work queued on it before
                                  txa_serve_submit, so a time loop may write the state on its own
                                  stream and submit without synchronising */
} txa_serve_request;

/* hbm_budget: bytes of the output arena (f and CRS values of admitted requests
 * until released); slots: requests in flight at once (1..32). */
int txa_serve_create(txa_ctx ctx, const int *mesh_ids, int n, int64_t hbm_budget, int slots, int *server_id);
/* Stops the worker after the running requests; queued ones fail. */
int txa_serve_destroy(txa_ctx ctx, int server_id);
/* Non-blocking: queues the request (TXA_ERR_NOMEM if it can never fit the budget). */
int txa_serve_submit(txa_ctx ctx, int server_id, const txa_serve_request *req, int64_t *request_id);
/* status: 0 queued, 1 running, 2 done (f / vals point at the result, device memory),
 * 3 failed (returns TXA_ERR_INVALID for a Dirichlet LID outside the DOF map,
 * TXA_ERR_DEVICE otherwise, with the message).  timeout_s > 0 waits up to that long
 * for the request to finish. */
int txa_serve_poll(txa_ctx ctx, int server_id, int64_t request_id, double timeout_s, int *status, const double **f,
                   const double **vals);
/* Returns the request's memory to the arena (done or failed requests only). */
int txa_serve_release(txa_ctx ctx, int server_id, int64_t request_id);
/* out[11]: completed, failed, p50 ms, p99 ms, mean ms, max ms (submit -> completion),
 * elements/s, requests/s (first submit -> last completion), peak requests in
 * flight, peak arena bytes, arena bytes in use */
int txa_serve_stats(txa_ctx ctx, int server_id, double *out);

#ifdef __cplusplus
}
#endif
#endif /* TXA_H */
