/*
 * oracle.c -- CPU restatement of Tianxin/Panzer's volume residual + Jacobian
 * assembly path.  TEST INFRASTRUCTURE ONLY: tests, the smoke check and the
 * benchmark's CPU-baseline leg call it as the checker.  Nothing in the product
 * (tianxin_amd/) links, loads or calls this file.
 *
 * Every stage follows the reference algorithm, cited by file:line
 * (paths relative to the reference repository root):
 *
 *   mesh numbering    adapters-stk/src/stk_interface/Panzer_STK_CubeHexMeshFactory.cpp:425-456
 *                     dof-mgr/test/cartesian_topology/CartesianConnManager.cpp:220-235 (6-tet split)
 *                     dof-mgr/test/cartesian_topology/CartesianConnManager.cpp:744-890 (edge/face/cell ids)
 *   GIDs (GUN, 1 rank) dof-mgr/src/Panzer_DOFManager.cpp:474-715, 717-865, 868-975, 1260-1349
 *   element DOF order dof-mgr/src/Panzer_FieldAggPattern.cpp:202-246
 *   owned_/LIDs       dof-mgr/src/Panzer_DOFManager.cpp:580-640; Panzer_GlobalIndexer.hpp:605-640
 *   CRS graph         disc-fe/src/lof/Panzer_TpetraLinearObjFactory_impl.hpp:559-650
 *   geometry          disc-fe/src/Panzer_IntegrationValues2.cpp:949-1054, 1059-1221
 *   basis values      disc-fe/src/Panzer_BasisValues2_impl.hpp:1037-1140, 1379-1460
 *   worksets          disc-fe/src/Panzer_Workset_Builder_impl.hpp:62-200 (cached per workset)
 *   gather + seeds    disc-fe/src/evaluators/Panzer_GatherSolution_Tpetra_impl.hpp:549-650
 *   DOF, DOFGradient  disc-fe/src/evaluators/Panzer_DOF_Functors.hpp:156-279;
 *                     disc-fe/src/evaluators/Panzer_DOFGradient_impl.hpp:60-160
 *   integrators       disc-fe/src/evaluators/Panzer_Integrator_GradBasisDotVector_impl.hpp:225-294
 *                     disc-fe/src/evaluators/Panzer_Integrator_BasisTimesScalar_impl.hpp:215-270
 *   Jacobian (FAD)    disc-fe/src/Panzer_Traits.hpp:95,106 (Sacado DFad, length = element DOFs)
 *   scatter           disc-fe/src/evaluators/Panzer_ScatterResidual_Tpetra_impl.hpp:375-441
 *   Dirichlet         disc-fe/src/lof/Panzer_TpetraLinearObjContainer.hpp:228-237, 300-318
 *
 * Elasticity and the Newmark/Rayleigh terms have no counterpart in the
 * reference; they are written in the same evaluator style (DOFGradient ->
 * stress -> GradBasisDotVector, BasisTimesScalar for inertia) and are pinned by
 * closed forms and invariants in the tests.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "oracle.h"

#define MAXB 27   /* basis functions per element */
#define MAXQ 27   /* integration points per element */
#define MAXD 81   /* derivative length (element DOFs) */
#define MAXV 8    /* geometry vertices per element */

/* ------------------------------------------------------------------------ */
/* Structured brick meshes                                                   */
/* ------------------------------------------------------------------------ */

/* Hex vertex offsets in shards order (CubeHexMeshFactory.cpp:448-455). */
static const int HEXV[8][3] = {{0,0,0},{1,0,0},{1,1,0},{0,1,0},{0,0,1},{1,0,1},{1,1,1},{0,1,1}};
static const int QUADV[4][2] = {{0,0},{1,0},{1,1},{0,1}};
/* Six tets per hex over hex-local vertices (CartesianConnManager.cpp:228-235). */
static const int TETS_IN_HEX[6][4] = {{0,2,3,6},{0,3,7,6},{0,7,4,6},{0,5,6,4},{1,5,6,0},{1,6,2,0}};
/* shards Hexahedron edges and faces (vertex lists). */
static const int HEX_EDGE[12][2] = {{0,1},{1,2},{2,3},{3,0},{4,5},{5,6},{6,7},{7,4},{0,4},{1,5},{2,6},{3,7}};
static const int HEX_FACE[6][4] = {{0,1,5,4},{1,2,6,5},{2,3,7,6},{0,4,7,3},{0,3,2,1},{4,5,6,7}};

/* out[0]=nelem out[1]=conn entries/elem out[2]=geometry vertices/elem
 * out[3]=number of vertex nodes out[4]=entity-id upper bound out[5]=dim */
int orc_mesh_sizes(int topo, int nx, int ny, int nz, int64_t *out)
{
  int64_t NX = nx, NY = ny, NZ = nz;
  if (nx < 1 || ny < 1) return -1;
  if (topo == ORC_QUAD4) {
    out[0] = NX * NY; out[1] = 4; out[2] = 4; out[3] = (NX + 1) * (NY + 1);
    out[4] = out[3]; out[5] = 2; return 0;
  }
  if (nz < 1) return -1;
  int64_t nn = (NX + 1) * (NY + 1) * (NZ + 1);
  int64_t ne = NX * (NY + 1) * (NZ + 1) + (NX + 1) * NY * (NZ + 1) + (NX + 1) * (NY + 1) * NZ;
  int64_t nf = NX * NY * (NZ + 1) + (NX * (NY + 1) + (NX + 1) * NY) * NZ;
  int64_t nc = NX * NY * NZ;
  out[3] = nn; out[5] = 3;
  if (topo == ORC_HEX8) { out[0] = nc; out[1] = 8; out[2] = 8; out[4] = nn; return 0; }
  if (topo == ORC_TET4) { out[0] = 6 * nc; out[1] = 4; out[2] = 4; out[4] = nn; return 0; }
  if (topo == ORC_HEX27) { out[0] = nc; out[1] = 27; out[2] = 8; out[4] = nn + ne + nf + nc; return 0; }
  return -1;
}

static int64_t node_id3(int64_t nx, int64_t ny, int64_t x, int64_t y, int64_t z)
{
  return x + (nx + 1) * (y + (ny + 1) * z);
}

/* Entity ids of a Cartesian brick: nodes, then edges, then faces, then cells.
 * Within a z level: x-edges and y-edges interleaved by row, then the vertical
 * edges up to the next level; faces likewise (xy faces of the level, then the
 * vertical faces by row: xz faces then yz faces).  This is the numbering of
 * CartesianConnManager.cpp:790-868, restated by entity kind. */
static int64_t edge_x(int64_t nx, int64_t ny, int64_t x, int64_t y, int64_t z)
{
  int64_t ks = (nx + 1) * ny + nx * (ny + 1) + (nx + 1) * (ny + 1);
  return x + y * (2 * nx + 1) + z * ks;
}
static int64_t edge_y(int64_t nx, int64_t ny, int64_t x, int64_t y, int64_t z)
{
  int64_t ks = (nx + 1) * ny + nx * (ny + 1) + (nx + 1) * (ny + 1);
  return nx + x + y * (2 * nx + 1) + z * ks;
}
static int64_t edge_z(int64_t nx, int64_t ny, int64_t x, int64_t y, int64_t z)
{
  int64_t ks = (nx + 1) * ny + nx * (ny + 1) + (nx + 1) * (ny + 1);
  return z * ks + (nx + 1) * ny + nx * (ny + 1) + x + y * (nx + 1);
}
static int64_t face_xy(int64_t nx, int64_t ny, int64_t x, int64_t y, int64_t z)
{
  int64_t ks = nx * ny + nx * (ny + 1) + (nx + 1) * ny;
  return x + y * nx + z * ks;
}
static int64_t face_xz(int64_t nx, int64_t ny, int64_t x, int64_t y, int64_t z)
{
  int64_t ks = nx * ny + nx * (ny + 1) + (nx + 1) * ny;
  return z * ks + nx * ny + y * (2 * nx + 1) + x;
}
static int64_t face_yz(int64_t nx, int64_t ny, int64_t x, int64_t y, int64_t z)
{
  int64_t ks = nx * ny + nx * (ny + 1) + (nx + 1) * ny;
  return z * ks + nx * ny + y * (2 * nx + 1) + nx + x;
}

/* conn: [nelem][nconn] entity ids; elem_vert: [nelem][nvert] vertex node ids. */
int orc_mesh_build(int topo, int nx, int ny, int nz, int64_t *conn, int64_t *elem_vert)
{
  int64_t sz[6];
  if (orc_mesh_sizes(topo, nx, ny, nz, sz)) return -1;
  int64_t NX = nx, NY = ny, NZ = nz;
  if (topo == ORC_QUAD4) {
    for (int64_t y = 0; y < NY; ++y)
      for (int64_t x = 0; x < NX; ++x) {
        int64_t e = x + NX * y;
        for (int v = 0; v < 4; ++v) {
          int64_t n = (x + QUADV[v][0]) + (NX + 1) * (y + QUADV[v][1]);
          conn[e * 4 + v] = n; elem_vert[e * 4 + v] = n;
        }
      }
    return 0;
  }
  int64_t nn = sz[3];
  int64_t ne = NX * (NY + 1) * (NZ + 1) + (NX + 1) * NY * (NZ + 1) + (NX + 1) * (NY + 1) * NZ;
  int64_t nf = NX * NY * (NZ + 1) + (NX * (NY + 1) + (NX + 1) * NY) * NZ;
  for (int64_t z = 0; z < NZ; ++z)
    for (int64_t y = 0; y < NY; ++y)
      for (int64_t x = 0; x < NX; ++x) {
        int64_t c = x + NX * (y + NY * z);
        int64_t hv[8];
        for (int v = 0; v < 8; ++v)
          hv[v] = node_id3(NX, NY, x + HEXV[v][0], y + HEXV[v][1], z + HEXV[v][2]);
        if (topo == ORC_HEX8) {
          for (int v = 0; v < 8; ++v) { conn[c * 8 + v] = hv[v]; elem_vert[c * 8 + v] = hv[v]; }
        } else if (topo == ORC_TET4) {
          for (int t = 0; t < 6; ++t)
            for (int v = 0; v < 4; ++v) {
              int64_t n = hv[TETS_IN_HEX[t][v]];
              conn[(6 * c + t) * 4 + v] = n; elem_vert[(6 * c + t) * 4 + v] = n;
            }
        } else { /* HEX27: vertices, 12 edges, 6 faces, cell (subcell order) */
          int64_t *cc = conn + c * 27;
          for (int v = 0; v < 8; ++v) { cc[v] = hv[v]; elem_vert[c * 8 + v] = hv[v]; }
          cc[8]  = nn + edge_x(NX, NY, x, y, z);
          cc[9]  = nn + edge_y(NX, NY, x + 1, y, z);
          cc[10] = nn + edge_x(NX, NY, x, y + 1, z);
          cc[11] = nn + edge_y(NX, NY, x, y, z);
          cc[12] = nn + edge_x(NX, NY, x, y, z + 1);
          cc[13] = nn + edge_y(NX, NY, x + 1, y, z + 1);
          cc[14] = nn + edge_x(NX, NY, x, y + 1, z + 1);
          cc[15] = nn + edge_y(NX, NY, x, y, z + 1);
          cc[16] = nn + edge_z(NX, NY, x, y, z);
          cc[17] = nn + edge_z(NX, NY, x + 1, y, z);
          cc[18] = nn + edge_z(NX, NY, x + 1, y + 1, z);
          cc[19] = nn + edge_z(NX, NY, x, y + 1, z);
          cc[20] = nn + ne + face_xz(NX, NY, x, y, z);      /* y = -1 */
          cc[21] = nn + ne + face_yz(NX, NY, x + 1, y, z);  /* x = +1 */
          cc[22] = nn + ne + face_xz(NX, NY, x, y + 1, z);  /* y = +1 */
          cc[23] = nn + ne + face_yz(NX, NY, x, y, z);      /* x = -1 */
          cc[24] = nn + ne + face_xy(NX, NY, x, y, z);      /* z = -1 */
          cc[25] = nn + ne + face_xy(NX, NY, x, y, z + 1);  /* z = +1 */
          cc[26] = nn + ne + nf + c;
        }
      }
  return 0;
}

/* ------------------------------------------------------------------------ */
/* DOF manager: GUN numbering on one rank + first-touch owned_ + LIDs        */
/* ------------------------------------------------------------------------ */

static int cmp_i64(const void *a, const void *b)
{
  int64_t x = *(const int64_t *)a, y = *(const int64_t *)b;
  return (x > y) - (x < y);
}
static int cmp_i32(const void *a, const void *b)
{
  int32_t x = *(const int32_t *)a, y = *(const int32_t *)b;
  return (x > y) - (x < y);
}

static int64_t lower_bound_i64(const int64_t *a, int64_t n, int64_t key)
{
  int64_t lo = 0, hi = n;
  while (lo < hi) { int64_t m = (lo + hi) / 2; if (a[m] < key) lo = m + 1; else hi = m; }
  return lo;
}

/*
 * Every connectivity entity carries one DOF per field (nodal, edge, face and
 * cell patterns of HGRAD C1/C2 all have one DOF per subcell).
 *   overlap map  = sorted unique entity ids          (DOFManager.cpp:1260-1290)
 *   GUN          = entity by entity, field by field   (DOFManager.cpp:826-845)
 *   element GIDs = per entity (subcell order), per field (DOFManager.cpp:1294-1349,
 *                  FieldAggPattern.cpp:202-246)
 *   owned_       = first touch over elements         (DOFManager.cpp:580-640)
 *   LID          = position in owned_                (GlobalIndexer.hpp:605-640)
 * elem_gid/elem_lid: [nelem][nconn*nfields]; owned: capacity nentities*nfields.
 * Returns the number of DOFs, or -1.
 */
int64_t orc_dofmap(int64_t nelem, int nconn, const int64_t *conn, int nfields,
                   int64_t *elem_gid, int64_t *owned, int32_t *elem_lid)
{
  int64_t ncon_tot = nelem * nconn;
  int64_t *ids = (int64_t *)malloc(sizeof(int64_t) * (ncon_tot > 0 ? ncon_tot : 1));
  if (!ids) return -1;
  memcpy(ids, conn, sizeof(int64_t) * ncon_tot);
  qsort(ids, ncon_tot, sizeof(int64_t), cmp_i64);
  int64_t nuniq = 0;
  for (int64_t i = 0; i < ncon_tot; ++i)
    if (i == 0 || ids[i] != ids[i - 1]) ids[nuniq++] = ids[i];
  /* tagged multivector: one DOF per (entity, field); GID offset is the
   * exclusive scan in overlap-map order, fields contiguous within an entity */
  int ndof = nconn * nfields;
  for (int64_t e = 0; e < nelem; ++e)
    for (int c = 0; c < nconn; ++c) {
      int64_t r = lower_bound_i64(ids, nuniq, conn[e * nconn + c]);
      for (int f = 0; f < nfields; ++f)
        elem_gid[e * ndof + c * nfields + f] = r * nfields + f;
    }
  int64_t ndofs = nuniq * nfields;
  free(ids);
  /* first-touch owned_ and LIDs: GIDs are dense in [0, ndofs) on one rank */
  int32_t *lid_of = (int32_t *)malloc(sizeof(int32_t) * (ndofs > 0 ? ndofs : 1));
  if (!lid_of) return -1;
  for (int64_t g = 0; g < ndofs; ++g) lid_of[g] = -1;
  int64_t next = 0;
  for (int64_t e = 0; e < nelem; ++e)
    for (int i = 0; i < ndof; ++i) {
      int64_t g = elem_gid[e * ndof + i];
      if (lid_of[g] < 0) { lid_of[g] = (int32_t)next; owned[next] = g; ++next; }
    }
  for (int64_t p = 0; p < nelem * ndof; ++p) elem_lid[p] = lid_of[elem_gid[p]];
  free(lid_of);
  return next == ndofs ? ndofs : -1;
}

/* ------------------------------------------------------------------------ */
/* CRS graph: union of element closures, rows = LIDs, columns sorted unique  */
/* (TpetraLinearObjFactory_impl.hpp:559-650 + Tpetra fillComplete)            */
/* ------------------------------------------------------------------------ */

static int64_t *graph_candidates(int64_t nelem, int ndof, const int32_t *elem_lid, int64_t nrows,
                                 int32_t **cand_out)
{
  int64_t *cptr = (int64_t *)calloc(nrows + 1, sizeof(int64_t));
  for (int64_t p = 0; p < nelem * ndof; ++p) cptr[elem_lid[p] + 1] += ndof;
  for (int64_t r = 0; r < nrows; ++r) cptr[r + 1] += cptr[r];
  int32_t *cand = (int32_t *)malloc(sizeof(int32_t) * (cptr[nrows] > 0 ? cptr[nrows] : 1));
  int64_t *cur = (int64_t *)malloc(sizeof(int64_t) * (nrows > 0 ? nrows : 1));
  memcpy(cur, cptr, sizeof(int64_t) * nrows);
  for (int64_t e = 0; e < nelem; ++e)
    for (int i = 0; i < ndof; ++i) {
      int32_t r = elem_lid[e * ndof + i];
      for (int j = 0; j < ndof; ++j) cand[cur[r]++] = elem_lid[e * ndof + j];
    }
  free(cur);
  for (int64_t r = 0; r < nrows; ++r)
    qsort(cand + cptr[r], cptr[r + 1] - cptr[r], sizeof(int32_t), cmp_i32);
  *cand_out = cand;
  return cptr;
}

/* Fills rowptr[nrows+1]; returns nnz. */
int64_t orc_graph_rowptr(int64_t nelem, int ndof, const int32_t *elem_lid, int64_t nrows, int64_t *rowptr)
{
  int32_t *cand;
  int64_t *cptr = graph_candidates(nelem, ndof, elem_lid, nrows, &cand);
  rowptr[0] = 0;
  for (int64_t r = 0; r < nrows; ++r) {
    int64_t u = 0;
    for (int64_t k = cptr[r]; k < cptr[r + 1]; ++k)
      if (k == cptr[r] || cand[k] != cand[k - 1]) ++u;
    rowptr[r + 1] = rowptr[r] + u;
  }
  free(cand); free(cptr);
  return rowptr[nrows];
}

int orc_graph_cols(int64_t nelem, int ndof, const int32_t *elem_lid, int64_t nrows, const int64_t *rowptr,
                   int32_t *col)
{
  int32_t *cand;
  int64_t *cptr = graph_candidates(nelem, ndof, elem_lid, nrows, &cand);
  for (int64_t r = 0; r < nrows; ++r) {
    int64_t o = rowptr[r];
    for (int64_t k = cptr[r]; k < cptr[r + 1]; ++k)
      if (k == cptr[r] || cand[k] != cand[k - 1]) col[o++] = cand[k];
    if (o != rowptr[r + 1]) { free(cand); free(cptr); return -1; }
  }
  free(cand); free(cptr);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Reference elements: bases and cubature                                    */
/* ------------------------------------------------------------------------ */

typedef struct {
  int topo, dim;
  int nb;   /* field basis functions */
  int nv;   /* geometry vertices */
  int nq;   /* integration points */
  double qp[MAXQ][3], qw[MAXQ];
  double N[MAXB][MAXQ], dN[MAXB][MAXQ][3];   /* field basis at IPs (reference) */
  double gdN[MAXV][MAXQ][3];                 /* geometry basis gradients at IPs */
} refelem;

static double lag2(int node, double t)  /* 1D quadratic Lagrange, nodes -1,0,1 */
{
  if (node < 0) return 0.5 * t * (t - 1.0);
  if (node == 0) return 1.0 - t * t;
  return 0.5 * t * (t + 1.0);
}
static double dlag2(int node, double t)
{
  if (node < 0) return t - 0.5;
  if (node == 0) return -2.0 * t;
  return t + 0.5;
}

/* Reference coordinates (-1,0,1 per axis) of the 27 Q2 DOF nodes in subcell
 * order: vertices, edge midpoints, face centres, cell centre. */
static void q2_nodes(int t[27][3])
{
  int vc[8][3];
  for (int v = 0; v < 8; ++v)
    for (int d = 0; d < 3; ++d) vc[v][d] = 2 * HEXV[v][d] - 1;
  for (int v = 0; v < 8; ++v)
    for (int d = 0; d < 3; ++d) t[v][d] = vc[v][d];
  for (int e = 0; e < 12; ++e)
    for (int d = 0; d < 3; ++d) t[8 + e][d] = (vc[HEX_EDGE[e][0]][d] + vc[HEX_EDGE[e][1]][d]) / 2;
  for (int f = 0; f < 6; ++f)
    for (int d = 0; d < 3; ++d) {
      int s = 0;
      for (int k = 0; k < 4; ++k) s += vc[HEX_FACE[f][k]][d];
      t[20 + f][d] = s / 4;
    }
  for (int d = 0; d < 3; ++d) t[26][d] = 0;
}

static void refelem_init(refelem *r, int topo)
{
  memset(r, 0, sizeof(*r));
  r->topo = topo;
  if (topo == ORC_QUAD4 || topo == ORC_HEX8 || topo == ORC_HEX27) {
    int dim = topo == ORC_QUAD4 ? 2 : 3;
    int ng = topo == ORC_HEX27 ? 3 : 2;
    double g1[3], w1[3];
    if (ng == 2) { g1[0] = -1.0 / sqrt(3.0); g1[1] = 1.0 / sqrt(3.0); w1[0] = w1[1] = 1.0; }
    else { g1[0] = -sqrt(0.6); g1[1] = 0.0; g1[2] = sqrt(0.6); w1[0] = w1[2] = 5.0 / 9.0; w1[1] = 8.0 / 9.0; }
    r->dim = dim;
    int q = 0;
    if (dim == 2) {
      for (int j = 0; j < ng; ++j)
        for (int i = 0; i < ng; ++i) { r->qp[q][0] = g1[i]; r->qp[q][1] = g1[j]; r->qw[q] = w1[i] * w1[j]; ++q; }
    } else {
      for (int k = 0; k < ng; ++k)
        for (int j = 0; j < ng; ++j)
          for (int i = 0; i < ng; ++i) {
            r->qp[q][0] = g1[i]; r->qp[q][1] = g1[j]; r->qp[q][2] = g1[k];
            r->qw[q] = w1[i] * w1[j] * w1[k]; ++q;
          }
    }
    r->nq = q;
    r->nv = dim == 2 ? 4 : 8;
    for (int qq = 0; qq < q; ++qq) {
      const double *p = r->qp[qq];
      for (int v = 0; v < r->nv; ++v) {
        double s[3];
        for (int d = 0; d < dim; ++d) s[d] = dim == 2 ? 2.0 * QUADV[v][d] - 1.0 : 2.0 * HEXV[v][d] - 1.0;
        double f[3], df[3];
        for (int d = 0; d < dim; ++d) { f[d] = 0.5 * (1.0 + s[d] * p[d]); df[d] = 0.5 * s[d]; }
        double val = 1.0;
        for (int d = 0; d < dim; ++d) val *= f[d];
        for (int d = 0; d < dim; ++d) {
          double g = df[d];
          for (int e = 0; e < dim; ++e) if (e != d) g *= f[e];
          r->gdN[v][qq][d] = g;
        }
        if (topo != ORC_HEX27) {
          r->N[v][qq] = val;
          for (int d = 0; d < dim; ++d) r->dN[v][qq][d] = r->gdN[v][qq][d];
        }
      }
    }
    if (topo == ORC_HEX27) {
      int t[27][3];
      q2_nodes(t);
      r->nb = 27;
      for (int qq = 0; qq < q; ++qq) {
        const double *p = r->qp[qq];
        for (int b = 0; b < 27; ++b) {
          double l[3], dl[3];
          for (int d = 0; d < 3; ++d) { l[d] = lag2(t[b][d], p[d]); dl[d] = dlag2(t[b][d], p[d]); }
          r->N[b][qq] = l[0] * l[1] * l[2];
          r->dN[b][qq][0] = dl[0] * l[1] * l[2];
          r->dN[b][qq][1] = l[0] * dl[1] * l[2];
          r->dN[b][qq][2] = l[0] * l[1] * dl[2];
        }
      }
    } else {
      r->nb = r->nv;
    }
    return;
  }
  if (topo == ORC_WEDGE6) {
    /* HGRAD_WEDGE_C1 on shards Wedge<6> (Panzer_IntrepidBasisFactory.hpp:159-184 selects the
     * Intrepid2 HGRAD basis of the cell topology): nodes (0,0,-1) (1,0,-1) (0,1,-1) and the same
     * triangle at zeta = +1; N_i = lam_i (1 - zeta)/2, N_{i+3} = lam_i (1 + zeta)/2 with
     * lam = (1 - xi - eta, xi, eta).  Cubature of degree 2 as a tensor rule (the reference's
     * IntegrationRule asks Intrepid2 for degree 2, Panzer_IntegrationRule.cpp:130-166): the
     * 3-point degree-2 triangle rule (1/6,1/6) (2/3,1/6) (1/6,2/3), weights 1/6, times 2-point
     * Gauss in zeta, triangle point fastest. */
    static const double tq[3][2] = {{1.0 / 6.0, 1.0 / 6.0}, {2.0 / 3.0, 1.0 / 6.0}, {1.0 / 6.0, 2.0 / 3.0}};
    const double zg[2] = {-1.0 / sqrt(3.0), 1.0 / sqrt(3.0)};
    r->dim = 3; r->nb = 6; r->nv = 6; r->nq = 6;
    for (int k = 0; k < 2; ++k)
      for (int t = 0; t < 3; ++t) {
        const int q = 3 * k + t;
        r->qp[q][0] = tq[t][0]; r->qp[q][1] = tq[t][1]; r->qp[q][2] = zg[k];
        r->qw[q] = 1.0 / 6.0;
      }
    for (int q = 0; q < 6; ++q) {
      const double x = r->qp[q][0], y = r->qp[q][1], z = r->qp[q][2];
      const double lam[3] = {1.0 - x - y, x, y};
      const double dlam[3][2] = {{-1.0, -1.0}, {1.0, 0.0}, {0.0, 1.0}};
      for (int l = 0; l < 2; ++l) {          /* l = 0: bottom (1 - zeta)/2, l = 1: top (1 + zeta)/2 */
        const double h = l == 0 ? 0.5 * (1.0 - z) : 0.5 * (1.0 + z), dh = l == 0 ? -0.5 : 0.5;
        for (int i = 0; i < 3; ++i) {
          const int v = 3 * l + i;
          r->N[v][q] = lam[i] * h;
          r->dN[v][q][0] = dlam[i][0] * h;
          r->dN[v][q][1] = dlam[i][1] * h;
          r->dN[v][q][2] = lam[i] * dh;
          for (int d = 0; d < 3; ++d) r->gdN[v][q][d] = r->dN[v][q][d];
        }
      }
    }
    return;
  }
  /* TET4: linear basis, degree-2 four-point rule */
  r->dim = 3; r->nb = 4; r->nv = 4; r->nq = 4;
  double a = 0.5854101966249685, b = 0.1381966011250105;
  for (int qq = 0; qq < 4; ++qq) {
    for (int d = 0; d < 3; ++d) r->qp[qq][d] = b;
    if (qq > 0) r->qp[qq][qq - 1] = a;
    r->qw[qq] = 1.0 / 24.0;
  }
  for (int qq = 0; qq < 4; ++qq) {
    const double *p = r->qp[qq];
    double n4[4] = {1.0 - p[0] - p[1] - p[2], p[0], p[1], p[2]};
    double g4[4][3] = {{-1, -1, -1}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    for (int v = 0; v < 4; ++v) {
      r->N[v][qq] = n4[v];
      for (int d = 0; d < 3; ++d) { r->dN[v][qq][d] = g4[v][d]; r->gdN[v][qq][d] = g4[v][d]; }
    }
  }
}

/* ------------------------------------------------------------------------ */
/* Forward-mode AD value (Sacado DFad<double> semantics, length = ndof)      */
/* ------------------------------------------------------------------------ */

typedef struct { double v; double dx[MAXD]; } fad;

static void fad_zero(fad *a, int n) { a->v = 0.0; memset(a->dx, 0, sizeof(double) * n); }
static void fad_axpy(fad *y, double s, const fad *x, int n)   /* y += s*x */
{
  y->v += s * x->v;
  for (int i = 0; i < n; ++i) y->dx[i] += s * x->dx[i];
}

/* ------------------------------------------------------------------------ */
/* Problem, cached worksets, evaluation                                      */
/* ------------------------------------------------------------------------ */

struct orc_ctx {
  orc_problem p;
  refelem re;
  int nf;                /* fields per basis function */
  int nws;
  /* cached per cell (IntegrationValues2 + BasisValues2 of every workset) */
  double *wm;            /* [nelem][nq]           weighted measure */
  double *gradN;         /* [nelem][nb][nq][dim]  grad_basis */
  double *wgradN;        /* [nelem][nb][nq][dim]  weighted_grad_basis */
  double *wN;            /* [nelem][nb][nq]       weighted_basis_scalar */
};

static void inv_det(int dim, const double J[3][3], double Ji[3][3], double *det)
{
  if (dim == 2) {
    double d = J[0][0] * J[1][1] - J[0][1] * J[1][0];
    Ji[0][0] = J[1][1] / d; Ji[0][1] = -J[0][1] / d;
    Ji[1][0] = -J[1][0] / d; Ji[1][1] = J[0][0] / d;
    *det = d; return;
  }
  double c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
  double c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
  double c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
  double d = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
  Ji[0][0] = c00 / d;
  Ji[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) / d;
  Ji[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) / d;
  Ji[1][0] = c01 / d;
  Ji[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) / d;
  Ji[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) / d;
  Ji[2][0] = c02 / d;
  Ji[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) / d;
  Ji[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) / d;
  *det = d;
}

/* Build the cached worksets: IntegrationValues2 (J, detJ, J^-1, weighted
 * measure) and BasisValues2 (grad_basis = J^-T grad_ref, weighted copies). */
orc_ctx *orc_setup(const orc_problem *p)
{
  orc_ctx *c = (orc_ctx *)calloc(1, sizeof(orc_ctx));
  if (!c) return NULL;
  c->p = *p;
  refelem_init(&c->re, p->topo);
  const refelem *r = &c->re;
  int dim = r->dim, nb = r->nb, nq = r->nq;
  c->nf = p->physics == ORC_ELASTICITY ? dim : 1;
  if (nb * c->nf != p->ndof || r->nv != p->nvert) { free(c); return NULL; }
  int64_t E = p->nelem;
  c->wm = (double *)malloc(sizeof(double) * E * nq);
  c->gradN = (double *)malloc(sizeof(double) * E * nb * nq * dim);
  c->wgradN = (double *)malloc(sizeof(double) * E * nb * nq * dim);
  c->wN = (double *)malloc(sizeof(double) * E * nb * nq);
  if (!c->wm || !c->gradN || !c->wgradN || !c->wN) { free(c->wm); free(c->gradN); free(c->wgradN); free(c->wN); free(c); return NULL; }
  int ws = p->workset_size > 0 ? p->workset_size : 1000;
  c->nws = (int)((E + ws - 1) / ws);
#ifdef _OPENMP
  if (p->nthreads > 0) omp_set_num_threads(p->nthreads);
#pragma omp parallel for schedule(dynamic)
#endif
  for (int w = 0; w < c->nws; ++w) {
    int64_t c0 = (int64_t)w * ws, c1 = c0 + ws < E ? c0 + ws : E;
    for (int64_t e = c0; e < c1; ++e) {
      double X[MAXV][3];
      for (int v = 0; v < r->nv; ++v)
        for (int d = 0; d < dim; ++d) X[v][d] = p->coords[p->elem_vert[e * r->nv + v] * dim + d];
      double sign = 1.0;
      for (int q = 0; q < nq; ++q) {
        double J[3][3] = {{0}}, Ji[3][3], det;
        /* CellTools::setJacobian: J_ij = sum_v X_v,i dN_v/dxi_j */
        for (int v = 0; v < r->nv; ++v)
          for (int i = 0; i < dim; ++i)
            for (int j = 0; j < dim; ++j) J[i][j] += X[v][i] * r->gdN[v][q][j];
        inv_det(dim, J, Ji, &det);
        /* computeCellMeasure: w*detJ, sign flipped when detJ at the first point is negative */
        if (q == 0) sign = det < 0.0 ? -1.0 : 1.0;
        double wmq = r->qw[q] * det * sign;
        c->wm[e * nq + q] = wmq;
        for (int b = 0; b < nb; ++b) {
          c->wN[(e * nb + b) * nq + q] = wmq * r->N[b][q];
          for (int i = 0; i < dim; ++i) {
            double g = 0.0;   /* HGRADtransformGRAD: (J^-T grad_ref)_i = sum_j Ji[j][i] dN_j */
            for (int j = 0; j < dim; ++j) g += Ji[j][i] * r->dN[b][q][j];
            c->gradN[((e * nb + b) * nq + q) * dim + i] = g;
            c->wgradN[((e * nb + b) * nq + q) * dim + i] = wmq * g;
          }
        }
      }
    }
  }
  return c;
}

void orc_free(orc_ctx *c)
{
  if (!c) return;
  free(c->wm); free(c->gradN); free(c->wgradN); free(c->wN); free(c);
}

static int64_t row_find(const int64_t *rowptr, const int32_t *col, int32_t row, int32_t key)
{
  int64_t lo = rowptr[row], hi = rowptr[row + 1];
  while (lo < hi) { int64_t m = (lo + hi) / 2; if (col[m] < key) lo = m + 1; else hi = m; }
  return (lo < rowptr[row + 1] && col[lo] == key) ? lo : -1;
}

static void atomic_add(double *p, double v)
{
#ifdef _OPENMP
#pragma omp atomic
#endif
  *p += v;
}

/* Scatter one element: residual atomic_add + sumIntoValues(row, elemLIDs,
 * derivative row, sorted, atomic)  (ScatterResidual_Tpetra_impl.hpp:375-413). */
static int scatter_cell(const orc_ctx *c, int64_t e, const double *Rv, const double *Jrows, int want_jac,
                        double *f, double *vals)
{
  int nd = c->p.ndof;
  const int32_t *lids = c->p.elem_lid + e * nd;
  for (int i = 0; i < nd; ++i) {
    int32_t row = lids[i];
    if (f) atomic_add(&f[row], Rv[i]);
    if (want_jac && vals)
      for (int j = 0; j < nd; ++j) {
        int64_t pos = row_find(c->p.rowptr, c->p.col, row, lids[j]);
        if (pos < 0) return -1;
        atomic_add(&vals[pos], Jrows[i * nd + j]);
      }
  }
  return 0;
}

/* Reference-style evaluation: gather with FAD seeds, DOF / DOFGradient,
 * closure (k or stress), integrators, scatter.  Residual-only if !want_jac
 * (then derivative arrays stay zero and are not scattered). */
static int eval_cell_fad(const orc_ctx *c, const orc_state *s, int64_t e, double *Rv, double *Jrows)
{
  const refelem *r = &c->re;
  const orc_problem *p = &c->p;
  int dim = r->dim, nb = r->nb, nq = r->nq, nf = c->nf, nd = p->ndof;
  int n = s->want_jac ? nd : 0;
  const int32_t *lids = p->elem_lid + e * nd;
  const double *wN = c->wN + e * nb * nq;
  const double *wG = c->wgradN + e * nb * nq * dim;
  const double *G = c->gradN + e * nb * nq * dim;
  static __thread fad U[MAXD], V[MAXD], A[MAXD], R[MAXD];
  static __thread fad gu[MAXQ][3][3], gv[MAXQ][3][3], uq[MAXQ][3], vq[MAXQ][3], aq[MAXQ][3];
  int tr = s->transient && s->xdot;
  /* GatherSolution_Tpetra<Jacobian>: value + seed dx[offset] */
  for (int i = 0; i < nd; ++i) {
    fad_zero(&U[i], n); U[i].v = s->x[lids[i]]; if (n) U[i].dx[i] = s->beta;
    if (tr) { fad_zero(&V[i], n); V[i].v = s->xdot[lids[i]]; if (n) V[i].dx[i] = s->alpha; }
    if (tr && s->xddot) { fad_zero(&A[i], n); A[i].v = s->xddot[lids[i]]; if (n) A[i].dx[i] = s->gamma2; }
  }
  /* DOF and DOFGradient at the integration points (field f of basis b sits at
   * element offset b*nf + f) */
  for (int q = 0; q < nq; ++q)
    for (int f = 0; f < nf; ++f) {
      fad_zero(&uq[q][f], n); fad_zero(&vq[q][f], n); fad_zero(&aq[q][f], n);
      for (int d = 0; d < dim; ++d) { fad_zero(&gu[q][f][d], n); fad_zero(&gv[q][f][d], n); }
      for (int b = 0; b < nb; ++b) {
        int o = b * nf + f;
        double Nbq = r->N[b][q];
        fad_axpy(&uq[q][f], Nbq, &U[o], n);
        if (tr) fad_axpy(&vq[q][f], Nbq, &V[o], n);
        if (tr && s->xddot) fad_axpy(&aq[q][f], Nbq, &A[o], n);
        for (int d = 0; d < dim; ++d) {
          double g = G[(b * nq + q) * dim + d];
          fad_axpy(&gu[q][f][d], g, &U[o], n);
          if (tr) fad_axpy(&gv[q][f][d], g, &V[o], n);
        }
      }
    }
  for (int i = 0; i < nd; ++i) fad_zero(&R[i], n);
  if (p->physics == ORC_DIFFUSION) {
    double k = p->coef[e];
    for (int q = 0; q < nq; ++q)
      for (int b = 0; b < nb; ++b) {
        /* Integrator_GradBasisDotVector: += w grad N_b . multiplier * flux * field multiplier(k) */
        for (int d = 0; d < dim; ++d) fad_axpy(&R[b], wG[(b * nq + q) * dim + d] * 1.0 * k, &gu[q][0][d], n);
        /* Integrator_BasisTimesScalar (transient): += wN_b * 1 * rho_c * udot */
        if (tr) fad_axpy(&R[b], wN[b * nq + q] * p->rho, &vq[q][0], n);
        /* Integrator_BasisTimesScalar (source): += wN_b * (-1) * f */
        R[b].v += wN[b * nq + q] * (-1.0) * p->src[0];
      }
  } else {
    double Ey = p->coef[e], nu = p->nu;
    double lam = Ey * nu / ((1.0 + nu) * (1.0 - 2.0 * nu)), mu = Ey / (2.0 * (1.0 + nu));
    static __thread fad sig[3][3], tmp;
    for (int q = 0; q < nq; ++q) {
      for (int pass = 0; pass < 2; ++pass) {
        if (pass == 1 && !(tr && p->rayleigh_b != 0.0)) break;
        fad (*g)[3] = pass == 0 ? gu[q] : gv[q];
        double scale = pass == 0 ? 1.0 : p->rayleigh_b;
        /* stress sigma_ij = lam tr(eps) delta_ij + mu (du_i/dx_j + du_j/dx_i) */
        fad_zero(&tmp, n);
        for (int d = 0; d < dim; ++d) fad_axpy(&tmp, 1.0, &g[d][d], n);
        for (int i = 0; i < dim; ++i)
          for (int j = 0; j < dim; ++j) {
            fad_zero(&sig[i][j], n);
            fad_axpy(&sig[i][j], mu, &g[i][j], n);
            fad_axpy(&sig[i][j], mu, &g[j][i], n);
            if (i == j) fad_axpy(&sig[i][j], lam, &tmp, n);
          }
        for (int b = 0; b < nb; ++b)
          for (int i = 0; i < dim; ++i)
            for (int d = 0; d < dim; ++d)
              fad_axpy(&R[b * nf + i], wG[(b * nq + q) * dim + d] * scale, &sig[i][d], n);
      }
      for (int b = 0; b < nb; ++b)
        for (int i = 0; i < dim; ++i) {
          double wn = wN[b * nq + q];
          if (tr) {
            if (s->xddot) fad_axpy(&R[b * nf + i], wn * p->rho, &aq[q][i], n);
            if (p->rayleigh_a != 0.0) fad_axpy(&R[b * nf + i], wn * p->rho * p->rayleigh_a, &vq[q][i], n);
          }
          R[b * nf + i].v += wn * (-1.0) * p->src[i];
        }
    }
  }
  for (int i = 0; i < nd; ++i) {
    Rv[i] = R[i].v;
    if (n) for (int j = 0; j < nd; ++j) Jrows[i * nd + j] = R[i].dx[j];
  }
  return 0;
}

/* Explicit tangent: K_e, M_e from the cached geometry; R = K u + C v + M a - F,
 * J = beta K + alpha C + gamma2 M with C = a_R M + b_R K (elasticity) or
 * C = rho_c M (diffusion, first order in time). */
static int eval_cell_explicit(const orc_ctx *c, const orc_state *s, int64_t e, double *Rv, double *Jrows)
{
  const refelem *r = &c->re;
  const orc_problem *p = &c->p;
  int dim = r->dim, nb = r->nb, nq = r->nq, nf = c->nf, nd = p->ndof;
  const int32_t *lids = p->elem_lid + e * nd;
  const double *wm = c->wm + e * nq;
  const double *G = c->gradN + e * nb * nq * dim;
  static __thread double K[MAXD * MAXD], M[MAXD * MAXD], F[MAXD];
  memset(K, 0, sizeof(double) * nd * nd);
  memset(M, 0, sizeof(double) * nd * nd);
  memset(F, 0, sizeof(double) * nd);
  int tr = s->transient && s->xdot;
  if (p->physics == ORC_DIFFUSION) {
    double k = p->coef[e];
    for (int q = 0; q < nq; ++q)
      for (int a = 0; a < nb; ++a) {
        F[a] += wm[q] * r->N[a][q] * p->src[0];
        for (int b = 0; b < nb; ++b) {
          double gg = 0.0;
          for (int d = 0; d < dim; ++d) gg += G[(a * nq + q) * dim + d] * G[(b * nq + q) * dim + d];
          K[a * nd + b] += wm[q] * k * gg;
          M[a * nd + b] += wm[q] * p->rho * r->N[a][q] * r->N[b][q];
        }
      }
  } else {
    double Ey = p->coef[e], nu = p->nu;
    double lam = Ey * nu / ((1.0 + nu) * (1.0 - 2.0 * nu)), mu = Ey / (2.0 * (1.0 + nu));
    for (int q = 0; q < nq; ++q)
      for (int a = 0; a < nb; ++a) {
        const double *ga = &G[(a * nq + q) * dim];
        for (int i = 0; i < dim; ++i) F[a * nf + i] += wm[q] * r->N[a][q] * p->src[i];
        for (int b = 0; b < nb; ++b) {
          const double *gb = &G[(b * nq + q) * dim];
          double gg = 0.0;
          for (int d = 0; d < dim; ++d) gg += ga[d] * gb[d];
          double m = wm[q] * p->rho * r->N[a][q] * r->N[b][q];
          for (int i = 0; i < dim; ++i) {
            M[(a * nf + i) * nd + b * nf + i] += m;
            for (int j = 0; j < dim; ++j)
              K[(a * nf + i) * nd + b * nf + j] +=
                  wm[q] * (lam * ga[i] * gb[j] + mu * ga[j] * gb[i] + (i == j ? mu * gg : 0.0));
          }
        }
      }
  }
  for (int i = 0; i < nd; ++i) {
    double acc = -F[i];
    for (int j = 0; j < nd; ++j) {
      double u = s->x[lids[j]];
      double Cij = 0.0;
      if (p->physics == ORC_DIFFUSION) Cij = M[i * nd + j];
      else Cij = p->rayleigh_a * M[i * nd + j] + p->rayleigh_b * K[i * nd + j];
      acc += K[i * nd + j] * u;
      if (tr) {
        acc += Cij * s->xdot[lids[j]];
        if (p->physics == ORC_ELASTICITY && s->xddot) acc += M[i * nd + j] * s->xddot[lids[j]];
      }
      if (s->want_jac) {
        double jv = s->beta * K[i * nd + j];
        if (tr) {
          jv += s->alpha * Cij;
          if (p->physics == ORC_ELASTICITY && s->xddot) jv += s->gamma2 * M[i * nd + j];
        }
        Jrows[i * nd + j] = jv;
      }
    }
    Rv[i] = acc;
  }
  return 0;
}

/* Volume assembly over all worksets into f (by LID) and the CRS values.
 * f and vals are accumulated into (the caller zeroes them, as
 * AssemblyEngine's beginFill/initializeContainer does).  Returns 0 or -1. */
int orc_evaluate(const orc_ctx *c, const orc_state *s, double *f, double *vals)
{
  if (!c || !s || !s->x) return -1;
  int ws = c->p.workset_size > 0 ? c->p.workset_size : 1000;
  int64_t E = c->p.nelem;
  int err = 0;
#ifdef _OPENMP
  if (c->p.nthreads > 0) omp_set_num_threads(c->p.nthreads);
#pragma omp parallel for schedule(dynamic) reduction(|:err)
#endif
  for (int w = 0; w < c->nws; ++w) {
    static __thread double Rv[MAXD], Jrows[MAXD * MAXD];
    int64_t c0 = (int64_t)w * ws, c1 = c0 + ws < E ? c0 + ws : E;
    for (int64_t e = c0; e < c1; ++e) {
      int rc = s->mode == ORC_MODE_FAD ? eval_cell_fad(c, s, e, Rv, Jrows)
                                       : eval_cell_explicit(c, s, e, Rv, Jrows);
      if (rc || scatter_cell(c, e, Rv, Jrows, s->want_jac, f, vals)) err |= 1;
    }
  }
  return err ? -1 : 0;
}

/* Dirichlet rows: f_i = x_i - g_i and the matrix row becomes the identity row
 * (TpetraLinearObjContainer.hpp:228-237, 300-318;
 * Tpetra::applyDirichletBoundaryConditionToLocalMatrixRows). */
int orc_dirichlet(int64_t n, const int32_t *lids, const double *g, const double *x, const int64_t *rowptr,
                  const int32_t *col, double *f, double *vals)
{
  for (int64_t i = 0; i < n; ++i) {
    int32_t r = lids[i];
    if (f && x) f[r] = x[r] - (g ? g[i] : 0.0);
    if (vals)
      for (int64_t k = rowptr[r]; k < rowptr[r + 1]; ++k) vals[k] = col[k] == r ? 1.0 : 0.0;
  }
  return 0;
}

/* Geometry access for tests: weighted measure and physical gradients of one cell. */
int orc_cell_geometry(const orc_ctx *c, int64_t e, double *wm, double *gradN)
{
  const refelem *r = &c->re;
  if (e < 0 || e >= c->p.nelem) return -1;
  memcpy(wm, c->wm + e * r->nq, sizeof(double) * r->nq);
  memcpy(gradN, c->gradN + e * r->nb * r->nq * r->dim, sizeof(double) * r->nb * r->nq * r->dim);
  return r->nq;
}

int orc_refelem_info(int topo, int32_t *out)
{
  refelem r;
  refelem_init(&r, topo);
  out[0] = r.dim; out[1] = r.nb; out[2] = r.nv; out[3] = r.nq;
  return 0;
}

/* Reference-element tables: N[nb][nq], dN[nb][nq][3], qp[nq][3], qw[nq]. */
int orc_refelem_tables(int topo, double *N, double *dN, double *qp, double *qw)
{
  refelem r;
  refelem_init(&r, topo);
  for (int b = 0; b < r.nb; ++b)
    for (int q = 0; q < r.nq; ++q) {
      N[b * r.nq + q] = r.N[b][q];
      for (int d = 0; d < 3; ++d) dN[(b * r.nq + q) * 3 + d] = r.dN[b][q][d];
    }
  for (int q = 0; q < r.nq; ++q) {
    for (int d = 0; d < 3; ++d) qp[q * 3 + d] = r.qp[q][d];
    qw[q] = r.qw[q];
  }
  return 0;
}
