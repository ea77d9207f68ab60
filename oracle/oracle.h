/* oracle.h -- declarations of the CPU oracle (test infrastructure only). */
#ifndef TXA_ORACLE_H
#define TXA_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif
enum { ORC_QUAD4 = 0, ORC_HEX8 = 1, ORC_TET4 = 2, ORC_HEX27 = 3, ORC_WEDGE6 = 4 };
enum { ORC_DIFFUSION = 0, ORC_ELASTICITY = 1 };
enum { ORC_MODE_FAD = 0, ORC_MODE_EXPLICIT = 1 };

typedef struct {
  int topo, physics;
  int64_t nelem;
  int nvert;
  const int64_t *elem_vert;
  const double *coords;
  int ndof;
  const int32_t *elem_lid;
  const double *coef;
  double nu, rho, src[3];
  double rayleigh_a, rayleigh_b;
  int workset_size;
  int nthreads;
  const int64_t *rowptr;
  const int32_t *col;
} orc_problem;

typedef struct {
  const double *x, *xdot, *xddot;
  double alpha, beta, gamma2;
  int transient;
  int mode;
  int want_jac;
} orc_state;

typedef struct orc_ctx orc_ctx;

int orc_mesh_sizes(int topo, int nx, int ny, int nz, int64_t *out);
int orc_mesh_build(int topo, int nx, int ny, int nz, int64_t *conn, int64_t *elem_vert);
int64_t orc_dofmap(int64_t nelem, int nconn, const int64_t *conn, int nfields, int64_t *elem_gid, int64_t *owned,
                   int32_t *elem_lid);
int64_t orc_graph_rowptr(int64_t nelem, int ndof, const int32_t *elem_lid, int64_t nrows, int64_t *rowptr);
int orc_graph_cols(int64_t nelem, int ndof, const int32_t *elem_lid, int64_t nrows, const int64_t *rowptr,
                   int32_t *col);
orc_ctx *orc_setup(const orc_problem *p);
void orc_free(orc_ctx *c);
int orc_evaluate(const orc_ctx *c, const orc_state *s, double *f, double *vals);
int orc_dirichlet(int64_t n, const int32_t *lids, const double *g, const double *x, const int64_t *rowptr,
                  const int32_t *col, double *f, double *vals);
#ifdef __cplusplus
}
#endif
#endif
