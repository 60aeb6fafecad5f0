/* A C caller of the ABI that lays its arrays out as an FBoxLib host does: per-box fabs,
 * column-major (x fastest), ghost layers around the valid box, for a 16^3 periodic grid split
 * into 8 boxes of 8^3 (a boxarray chopped at max_grid_size 8).  It runs apply_matrix and gmres
 * on those fabs and on one fab covering the domain, and checks that the two agree bit for bit,
 * that no ghost cell of an output was written and that a face two boxes hold is written to
 * both.  Exit status 0 on success.  Built by tests/test_gpu_c_harness.py with the host C
 * compiler against liblmfhd.so. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/lmfhd.h"

#define N 16
#define B 8  /* box edge */
#define NB 8
#define NG 2
#define SENT 1.0e300

static int fail(const char* what) {
  fprintf(stderr, "fab_harness: %s (%s)\n", what, lmfhd_last_error());
  return 1;
}

/* global packed field of extents e (x fastest) */
static double* field(const int e[3], unsigned seed) {
  double* f = malloc(sizeof(double) * e[0] * e[1] * e[2]);
  for (long q = 0; q < (long)e[0] * e[1] * e[2]; ++q) {
    seed = seed * 1103515245u + 12345u;
    f[q] = (double)((seed >> 8) & 0xffff) / 65536.0 - 0.5;
  }
  return f;
}

/* fab of box b (cell bounds lo..lo+B-1) with nodal flags nod, copied from global g (extents e) */
static lmfhd_fab box_fab(const int lo[3], const int nod[3], const double* g, const int e[3]) {
  lmfhd_fab f;
  int ext[3], s[3];
  for (int d = 0; d < 3; ++d) {
    f.lo[d] = lo[d];
    f.hi[d] = lo[d] + B - 1 + nod[d];
    ext[d] = B + nod[d];
    s[d] = ext[d] + 2 * NG;
  }
  f.ng = NG;
  f.data = malloc(sizeof(double) * s[0] * s[1] * s[2]);
  for (long q = 0; q < (long)s[0] * s[1] * s[2]; ++q) f.data[q] = SENT;
  if (g)
    for (int k = 0; k < ext[2]; ++k)
      for (int j = 0; j < ext[1]; ++j)
        for (int i = 0; i < ext[0]; ++i)
          f.data[(i + NG) + s[0] * ((j + NG) + s[1] * (long)(k + NG))] =
              g[(lo[0] + i) + e[0] * ((lo[1] + j) + e[1] * (long)(lo[2] + k))];
  return f;
}

/* single fab covering the domain (no ghosts) */
static lmfhd_fab dom_fab(double* g, const int e[3]) {
  lmfhd_fab f;
  f.data = g;
  for (int d = 0; d < 3; ++d) {
    f.lo[d] = 0;
    f.hi[d] = e[d] - 1;
  }
  f.ng = 0;
  return f;
}

/* box fab f equals global g on its valid region and holds SENT in its ghost cells */
static int check_box(const lmfhd_fab* f, const int nod[3], const double* g, const int e[3]) {
  int ext[3], s[3];
  for (int d = 0; d < 3; ++d) {
    ext[d] = B + nod[d];
    s[d] = ext[d] + 2 * NG;
  }
  for (int k = 0; k < s[2]; ++k)
    for (int j = 0; j < s[1]; ++j)
      for (int i = 0; i < s[0]; ++i) {
        const double v = f->data[i + s[0] * (j + s[1] * (long)k)];
        const int I = i - NG, J = j - NG, K = k - NG;
        const int valid = I >= 0 && I < ext[0] && J >= 0 && J < ext[1] && K >= 0 && K < ext[2];
        if (valid) {
          const double w = g[(f->lo[0] + I) + e[0] * ((f->lo[1] + J) + e[1] * (long)(f->lo[2] + K))];
          if (memcmp(&v, &w, sizeof v) != 0) return 0;
        } else if (v != SENT) {
          return 0;
        }
      }
  return 1;
}

int main(void) {
  lmfhd_desc d;
  lmfhd_desc_defaults(&d);
  d.dm = 3;
  d.n_cells[0] = d.n_cells[1] = d.n_cells[2] = N;
  d.dx = 1.0 / N;
  d.visc_type = -1;
  d.gmres_max_inner = 30;
  int blo[3 * NB], bhi[3 * NB], lo[NB][3];
  for (int b = 0; b < NB; ++b)
    for (int dd = 0; dd < 3; ++dd) {
      lo[b][dd] = ((b >> dd) & 1) * B;
      blo[3 * b + dd] = lo[b][dd];
      bhi[3 * b + dd] = lo[b][dd] + B - 1;
    }
  lmfhd_ctx *one = NULL, *mb = NULL;
  if (lmfhd_ctx_create(&d, &one) != LMFHD_OK) return fail("create (one box)");
  d.nboxes = NB;
  d.box_lo = blo;
  d.box_hi = bhi;
  if (lmfhd_ctx_create(&d, &mb) != LMFHD_OK) return fail("create (8 boxes)");

  int ef[3][3], ec[3] = {N, N, N}, ee[3][3];
  int nf[3][3], nc[3] = {0, 0, 0}, ne[3][3];
  for (int a = 0; a < 3; ++a)
    for (int dd = 0; dd < 3; ++dd) {
      nf[a][dd] = dd == a;
      ef[a][dd] = N + nf[a][dd];
      ne[a][dd] = a == 0 ? dd < 2 : (a == 1 ? dd != 1 : dd > 0); /* edges xy, xz, yz */
      ee[a][dd] = N + ne[a][dd];
    }
  /* coefficients: alpha = 1 + small, beta = 1 + small, beta_ed = 1 + small, gamma = 0 */
  double *alpha[3], *bed[3], *beta = field(ec, 7), *gam = calloc(N * N * N, sizeof(double));
  for (long q = 0; q < N * N * N; ++q) beta[q] = 1.0 + 0.5 * beta[q];
  for (int a = 0; a < 3; ++a) {
    alpha[a] = field(ef[a], 11 + a);
    bed[a] = field(ee[a], 21 + a);
    for (long q = 0; q < (long)ef[a][0] * ef[a][1] * ef[a][2]; ++q) alpha[a][q] = 1.0 + 0.5 * alpha[a][q];
    for (long q = 0; q < (long)ee[a][0] * ee[a][1] * ee[a][2]; ++q) bed[a][q] = 1.0 + 0.5 * bed[a][q];
  }
  lmfhd_fab A1[3], E1[3], B1 = dom_fab(beta, ec), G1 = dom_fab(gam, ec);
  lmfhd_fab AB[3 * NB], EB[3 * NB], BB[NB], GB[NB];
  for (int a = 0; a < 3; ++a) {
    A1[a] = dom_fab(alpha[a], ef[a]);
    E1[a] = dom_fab(bed[a], ee[a]);
    for (int b = 0; b < NB; ++b) {
      AB[a * NB + b] = box_fab(lo[b], nf[a], alpha[a], ef[a]);
      EB[a * NB + b] = box_fab(lo[b], ne[a], bed[a], ee[a]);
    }
  }
  for (int b = 0; b < NB; ++b) {
    BB[b] = box_fab(lo[b], nc, beta, ec);
    GB[b] = box_fab(lo[b], nc, gam, ec);
  }
  if (lmfhd_set_coeffs(one, A1, &B1, E1, &G1, 20.0) != LMFHD_OK) return fail("set_coeffs (one box)");
  if (lmfhd_set_coeffs(mb, AB, BB, EB, GB, 20.0) != LMFHD_OK) return fail("set_coeffs (8 boxes)");

  /* right-hand side: random faces (the duplicated periodic faces equal), zero pressure */
  double *bu[3], *bp = calloc(N * N * N, sizeof(double));
  for (int a = 0; a < 3; ++a) {
    bu[a] = field(ef[a], 31 + a);
    for (int k = 0; k < ef[a][2]; ++k)
      for (int j = 0; j < ef[a][1]; ++j)
        for (int i = 0; i < ef[a][0]; ++i) {
          const int I[3] = {i, j, k};
          if (I[a] == N) {
            int J[3] = {i, j, k};
            J[a] = 0;
            bu[a][i + ef[a][0] * (j + ef[a][1] * (long)k)] = bu[a][J[0] + ef[a][0] * (J[1] + ef[a][1] * (long)J[2])];
          }
        }
  }
  lmfhd_fab U1[3], UB[3 * NB], P1 = dom_fab(bp, ec), PB[NB];
  for (int a = 0; a < 3; ++a) {
    U1[a] = dom_fab(bu[a], ef[a]);
    for (int b = 0; b < NB; ++b) UB[a * NB + b] = box_fab(lo[b], nf[a], bu[a], ef[a]);
  }
  for (int b = 0; b < NB; ++b) PB[b] = box_fab(lo[b], nc, bp, ec);

  /* apply_matrix */
  double *ou[3], *op = malloc(sizeof(double) * N * N * N);
  lmfhd_fab O1[3], OB[3 * NB], OP1 = dom_fab(op, ec), OPB[NB];
  for (int a = 0; a < 3; ++a) {
    ou[a] = malloc(sizeof(double) * ef[a][0] * ef[a][1] * ef[a][2]);
    O1[a] = dom_fab(ou[a], ef[a]);
    for (int b = 0; b < NB; ++b) OB[a * NB + b] = box_fab(lo[b], nf[a], NULL, ef[a]);
  }
  for (int b = 0; b < NB; ++b) OPB[b] = box_fab(lo[b], nc, NULL, ec);
  if (lmfhd_apply_matrix(one, U1, &P1, O1, &OP1) != LMFHD_OK) return fail("apply_matrix (one box)");
  if (lmfhd_apply_matrix(mb, UB, PB, OB, OPB) != LMFHD_OK) return fail("apply_matrix (8 boxes)");
  for (int b = 0; b < NB; ++b) {
    for (int a = 0; a < 3; ++a)
      if (!check_box(&OB[a * NB + b], nf[a], ou[a], ef[a])) return fail("apply_matrix: face box differs");
    if (!check_box(&OPB[b], nc, op, ec)) return fail("apply_matrix: cell box differs");
  }

  /* gmres from x = 0 */
  double *xu[3], *xp = calloc(N * N * N, sizeof(double));
  lmfhd_fab X1[3], XB[3 * NB], XP1 = dom_fab(xp, ec), XPB[NB];
  for (int a = 0; a < 3; ++a) {
    xu[a] = calloc((size_t)ef[a][0] * ef[a][1] * ef[a][2], sizeof(double));
    X1[a] = dom_fab(xu[a], ef[a]);
    for (int b = 0; b < NB; ++b) XB[a * NB + b] = box_fab(lo[b], nf[a], xu[a], ef[a]);
  }
  for (int b = 0; b < NB; ++b) XPB[b] = box_fab(lo[b], nc, xp, ec);
  lmfhd_stats s1, s8;
  double n1, n8;
  if (lmfhd_gmres(one, U1, &P1, X1, &XP1, &n1, &s1) != LMFHD_OK) return fail("gmres (one box)");
  if (lmfhd_gmres(mb, UB, PB, XB, XPB, &n8, &s8) != LMFHD_OK) return fail("gmres (8 boxes)");
  if (s1.iters != s8.iters || !s1.converged || n1 != n8) return fail("gmres statistics differ");
  for (int b = 0; b < NB; ++b) {
    for (int a = 0; a < 3; ++a)
      if (!check_box(&XB[a * NB + b], nf[a], xu[a], ef[a])) return fail("gmres: face box differs");
    if (!check_box(&XPB[b], nc, xp, ec)) return fail("gmres: cell box differs");
  }
  printf("fab_harness: 8 boxes == one box (apply_matrix, gmres %d iterations)\n", s8.iters);
  lmfhd_ctx_destroy(one);
  lmfhd_ctx_destroy(mb);
  return 0;
}
