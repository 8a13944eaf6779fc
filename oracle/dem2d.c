/*
 * FP64 CPU oracle for the 2-D bonded-disk ice solver (test infrastructure only).
 *
 * This file restates, in plain C, the per-step force-and-integrate loop of
 * the reference solver `IceDEM` (src/tensordem/dem.py).  It is the checker the
 * HIP kernels are compared against; it is never linked into the product.
 *
 * What differs from the reference, and why the results are still the same:
 *   - The reference evaluates all N(N-1)/2 pairs every call (dem.py:116-118,
 *     238-292).  Only alive bonds and overlapping non-alive pairs carry force,
 *     so this file visits the bond list plus the pairs a uniform grid of cell
 *     size 2r reports as overlapping.  Pairs that carry no force contribute
 *     +-0.0 in the reference, which leaves every sum unchanged.
 *   - Forces are accumulated per particle in the order `index_add_` uses in
 *     dem.py:290-292: -drag*v, then +f of every pair (p, q>p) by ascending q,
 *     then -f of every pair (q<p, p) by ascending q.
 *   - `torch.linalg.pinv` of the 2x2 neighbourhood covariance (dem.py:202) is
 *     replaced by the closed-form inverse (rank-1 / rank-0 pseudo-inverse when
 *     the determinant is zero).  The two agree to a few ulps.
 *   - Reductions for the energy ledger are sequential sums in pair / particle
 *     order; torch uses blocked sums, so ledger channels agree to round-off.
 *
 * Every expression keeps the operand order of the Python source so that IEEE
 * rounding matches; build with -ffp-contract=off (no fused multiply-add).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* parameter block layout (shared with oracle/plate.py) */
enum {
  P_RADIUS = 0, P_KB, P_KC, P_EPS_T, P_EPS_S, P_CDAMP, P_DRAG, P_TOOL_R,
  P_DT, P_MASS, P_TOOL_X, P_TOOL_Y, P_TOOL_VX, P_TOOL_VY, P_TIME, P_NPARAM
};

/* accounting layout */
enum {
  A_EXT_WORK = 0, A_RELEASE, A_ACQUIRED, A_TOOL_WORK, A_DAMPING,
  A_EXT_IMP_X, A_EXT_IMP_Y, A_SUP_IMP_X, A_SUP_IMP_Y, A_DRAG_IMP_X, A_DRAG_IMP_Y,
  A_TOOL_IMP_X, A_TOOL_IMP_Y, A_NACC
};

/* last-step layout */
enum {
  L_T0 = 0, L_T1, L_WORK, L_SUP_X, L_SUP_Y, L_REACT_X, L_REACT_Y, L_EXT_X, L_EXT_Y,
  L_MOMENT, L_NLAST
};

/* forces() scalar outputs */
enum { S_RELEASED = 0, S_ACQUIRED, S_DAMP_POWER, S_TOOL_POWER, S_NEW_BREAKS, S_NSCAL };

static const double EPS64 = 2.220446049250313e-16; /* torch.finfo(float64).eps */

/* dem.py:182-208.  covariance / cross are accumulated with index_add_ over
 * the bond list: first every bond at its i end, then every bond at its j end. */
void orc_shear_strain(int n, const double *x, const double *x0, int nb,
                      const int32_t *bi, const int32_t *bj, double *s)
{
  double *cov = (double *)calloc((size_t)n * 4, sizeof(double));
  double *crs = (double *)calloc((size_t)n * 4, sizeof(double));
  for (int side = 0; side < 2; ++side) {
    for (int b = 0; b < nb; ++b) {
      int i = bi[b], j = bj[b];
      int p = side == 0 ? i : j;
      double rx = x0[2 * j] - x0[2 * i], ry = x0[2 * j + 1] - x0[2 * i + 1];
      cov[4 * p + 0] += rx * rx;
      cov[4 * p + 1] += rx * ry;
      cov[4 * p + 2] += ry * rx;
      cov[4 * p + 3] += ry * ry;
    }
  }
  for (int side = 0; side < 2; ++side) {
    for (int b = 0; b < nb; ++b) {
      int i = bi[b], j = bj[b];
      int p = side == 0 ? i : j;
      double rx = x0[2 * j] - x0[2 * i], ry = x0[2 * j + 1] - x0[2 * i + 1];
      double qx = x[2 * j] - x[2 * i], qy = x[2 * j + 1] - x[2 * i + 1];
      crs[4 * p + 0] += qx * rx;
      crs[4 * p + 1] += qx * ry;
      crs[4 * p + 2] += qy * rx;
      crs[4 * p + 3] += qy * ry;
    }
  }
  for (int p = 0; p < n; ++p) {
    double a = cov[4 * p], b = cov[4 * p + 1], c = cov[4 * p + 2], d = cov[4 * p + 3];
    double det = a * d - b * c;
    double i00, i01, i10, i11;
    if (det != 0.0) {
      i00 = d / det; i01 = -b / det; i10 = -c / det; i11 = a / det;
    } else {
      /* rank <= 1: pinv(C) = C / trace(C)^2 for a symmetric PSD C */
      double tr = a + d;
      if (tr > 0.0) {
        double t2 = tr * tr;
        i00 = a / t2; i01 = b / t2; i10 = c / t2; i11 = d / t2;
      } else {
        i00 = i01 = i10 = i11 = 0.0;
      }
    }
    const double *X = crs + 4 * p;
    double f00 = X[0] * i00 + X[1] * i10;
    double f01 = X[0] * i01 + X[1] * i11;
    double f10 = X[2] * i00 + X[3] * i10;
    double f11 = X[2] * i01 + X[3] * i11;
    /* (F^T F)_ab = F_0a F_0b + F_1a F_1b */
    double g00 = f00 * f00 + f10 * f10;
    double g01 = f00 * f01 + f10 * f11;
    double g11 = f01 * f01 + f11 * f11;
    double exx = 0.5 * (g00 - 1.0), exy = 0.5 * (g01 - 0.0), eyy = 0.5 * (g11 - 1.0);
    double dxy = exx - eyy;
    s[p] = sqrt(dxy * dxy + 4.0 * (exy * exy));
  }
  free(cov);
  free(crs);
}

/* ---------------- uniform-grid search for overlapping unbonded pairs ---------- */

typedef struct { int64_t key; int32_t idx; } keyed;

static int cmp_keyed(const void *a, const void *b)
{
  const keyed *u = (const keyed *)a, *v = (const keyed *)b;
  if (u->key != v->key) return u->key < v->key ? -1 : 1;
  return u->idx - v->idx;
}

static int64_t cell_key(int64_t cx, int64_t cy) { return cx * 4294967296LL + cy; }

typedef struct { int32_t i, j; } pair_t;

static int cmp_pair(const void *a, const void *b)
{
  const pair_t *u = (const pair_t *)a, *v = (const pair_t *)b;
  if (u->i != v->i) return u->i - v->i;
  return u->j - v->j;
}

/* sorted-partner lookup: is (i, j), i<j, a bond? */
static int is_bond(const int32_t *rowptr, const int32_t *col, int i, int j)
{
  int lo = rowptr[i], hi = rowptr[i + 1];
  while (lo < hi) {
    int mid = (lo + hi) >> 1;
    if (col[mid] < j) lo = mid + 1; else hi = mid;
  }
  return lo < rowptr[i + 1] && col[lo] == j;
}

/* All unordered pairs (i<j) with |x_j - x_i| < cutoff that are not bonds
 * (bonds are visited separately).  Returns a malloc'd sorted array.  Cells
 * are keyed by (floor(x/h), floor(y/h)) with h >= cutoff, so only the 3x3
 * block of cells around a particle can hold a partner. */
static pair_t *grid_pairs(int n, const double *x, double cutoff, const int32_t *rowptr,
                          const int32_t *col, int *count)
{
  double h = cutoff;
  keyed *k = (keyed *)malloc(sizeof(keyed) * (size_t)(n > 0 ? n : 1));
  int64_t *cx = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
  int64_t *cy = (int64_t *)malloc(sizeof(int64_t) * (size_t)(n > 0 ? n : 1));
  int m = 0;
  for (int p = 0; p < n; ++p) {
    double px = x[2 * p], py = x[2 * p + 1];
    if (!isfinite(px) || !isfinite(py) || fabs(px / h) > 1e9 || fabs(py / h) > 1e9) {
      cx[p] = INT64_MIN;
      continue;
    }
    cx[p] = (int64_t)floor(px / h);
    cy[p] = (int64_t)floor(py / h);
    k[m].key = cell_key(cx[p], cy[p]);
    k[m].idx = p;
    ++m;
  }
  qsort(k, (size_t)m, sizeof(keyed), cmp_keyed);
  int cap = 64, cnt = 0;
  pair_t *out = (pair_t *)malloc(sizeof(pair_t) * (size_t)cap);
  for (int p = 0; p < n; ++p) {
    if (cx[p] == INT64_MIN) continue;
    for (int ox = -1; ox <= 1; ++ox) {
      for (int oy = -1; oy <= 1; ++oy) {
        int64_t key = cell_key(cx[p] + ox, cy[p] + oy);
        int lo = 0, hi = m;
        while (lo < hi) {
          int mid = (lo + hi) >> 1;
          if (k[mid].key < key) lo = mid + 1; else hi = mid;
        }
        for (int t = lo; t < m && k[t].key == key; ++t) {
          int q = k[t].idx;
          if (q <= p) continue;
          double dx = x[2 * q] - x[2 * p], dy = x[2 * q + 1] - x[2 * p + 1];
          double d = sqrt(dx * dx + dy * dy);
          if (!(d < cutoff)) continue;
          if (is_bond(rowptr, col, p, q)) continue;
          if (cnt == cap) { cap *= 2; out = (pair_t *)realloc(out, sizeof(pair_t) * (size_t)cap); }
          out[cnt].i = p; out[cnt].j = q; ++cnt;
        }
      }
    }
  }
  qsort(out, (size_t)cnt, sizeof(pair_t), cmp_pair);
  free(k); free(cx); free(cy);
  *count = cnt;
  return out;
}

/* Build a CSR of partners q>p for every bond (i<j) so is_bond() works. */
static void bond_csr(int n, int nb, const int32_t *bi, const int32_t *bj, int32_t **rowptr_out,
                     int32_t **col_out)
{
  int32_t *rowptr = (int32_t *)calloc((size_t)n + 1, sizeof(int32_t));
  int32_t *col = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nb > 0 ? nb : 1));
  for (int b = 0; b < nb; ++b) rowptr[bi[b] + 1]++;
  for (int p = 0; p < n; ++p) rowptr[p + 1] += rowptr[p];
  int32_t *fill = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
  memcpy(fill, rowptr, sizeof(int32_t) * (size_t)n);
  for (int b = 0; b < nb; ++b) col[fill[bi[b]]++] = bj[b];
  /* bonds are given in triu order, so each row is already ascending */
  free(fill);
  *rowptr_out = rowptr;
  *col_out = col;
}

/* one force-carrying pair */
typedef struct { int32_t i, j; double fx, fy; } pforce;

typedef struct { int32_t partner; int32_t group; double fx, fy; } contrib;

static int cmp_contrib(const void *a, const void *b)
{
  const contrib *u = (const contrib *)a, *v = (const contrib *)b;
  if (u->group != v->group) return u->group - v->group;
  return u->partner - v->partner;
}

/* Polyline hull: hull_contact.py:7-75 for one particle.  Returns the force on
 * the particle for the given damping. */
static void hull_particle(const double *p, const double *v, const double *vert, int m,
                          const double *hv, double r_p, double r_h, double k, double c,
                          double *fx, double *fy)
{
  int best = 0;
  double best_d2 = INFINITY, bdx = 0, bdy = 0;
  for (int s = 0; s < m - 1; ++s) {
    double sx = vert[2 * s], sy = vert[2 * s + 1];
    double gx = vert[2 * s + 2] - sx, gy = vert[2 * s + 3] - sy;
    double len2 = gx * gx + gy * gy;
    double rx = p[0] - sx, ry = p[1] - sy;
    double fr = (rx * gx + ry * gy) / len2;
    if (fr < 0.0) fr = 0.0;
    if (fr > 1.0) fr = 1.0;
    double cx = sx + fr * gx, cy = sy + fr * gy;
    double dx = p[0] - cx, dy = p[1] - cy;
    double d2 = dx * dx + dy * dy;
    if (d2 < best_d2) { best_d2 = d2; best = s; bdx = dx; bdy = dy; }
  }
  double distance = sqrt(best_d2);
  double gx = vert[2 * best + 2] - vert[2 * best], gy = vert[2 * best + 3] - vert[2 * best + 1];
  double len = sqrt(gx * gx + gy * gy);
  double tx = gx / len, ty = gy / len;
  if (tx < 0 || (tx == 0 && ty < 0)) { tx = -tx; ty = -ty; }
  double fbx = -ty, fby = tx;
  double rvx = v[0] - hv[0], rvy = v[1] - hv[1];
  double approach = rvx * fbx + rvy * fby;
  if (approach > 0) { fbx = -fbx; fby = -fby; }
  double nx, ny;
  if (distance > EPS64) {
    double dd = distance > EPS64 ? distance : EPS64;
    nx = bdx / dd; ny = bdy / dd;
  } else {
    nx = fbx; ny = fby;
  }
  double ov = r_p + r_h - distance;
  if (ov < 0) ov = 0;
  double sp = rvx * nx + rvy * ny;
  double mag = k * ov - c * sp;
  if (mag < 0) mag = 0;
  if (!(ov > 0)) mag = 0;
  *fx = mag * nx;
  *fy = mag * ny;
}

/*
 * IceDEM.forces (dem.py:226-345).  Commits fracture into alive/failure_*,
 * writes force (n,2), reaction (2), and scal[S_*].
 *   hull == NULL: circular tool at (P_TOOL_X, P_TOOL_Y) moving with P_TOOL_V*.
 *   hull != NULL: polyline vertices (already in global position), hull
 *                 velocity in P_TOOL_V*.
 */
static void plate_forces(const double *prm, int n, const double *x, const double *v,
                         const double *x0, int nb, const int32_t *bi, const int32_t *bj,
                         const double *rest, uint8_t *alive, uint8_t *fmode, double *ftime,
                         double *fext, double *fenergy, const double *loads, const double *hull,
                         int hull_m, double *force, double *reaction, double *scal,
                         const double *eps_t_b, const double *eps_s_b)
{
  const double r = prm[P_RADIUS], kb = prm[P_KB], kc = prm[P_KC];
  const double cdamp = prm[P_CDAMP], drag = prm[P_DRAG];
  const double two_r = 2 * r;
  double *s = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
  orc_shear_strain(n, x, x0, nb, bi, bj, s);

  /* fracture commit (dem.py:238-277) */
  double released = 0.0, acquired = 0.0;
  int new_breaks = 0;
  double *dist_b = (double *)malloc(sizeof(double) * (size_t)(nb > 0 ? nb : 1));
  double *nx_b = (double *)malloc(sizeof(double) * (size_t)(nb > 0 ? nb : 1));
  double *ny_b = (double *)malloc(sizeof(double) * (size_t)(nb > 0 ? nb : 1));
  for (int b = 0; b < nb; ++b) {
    int i = bi[b], j = bj[b];
    double dx = x[2 * j] - x[2 * i], dy = x[2 * j + 1] - x[2 * i + 1];
    double d = sqrt(dx * dx + dy * dy);
    double nx, ny;
    if (d > 0) {
      double dd = d > EPS64 ? d : EPS64;
      nx = dx / dd; ny = dy / dd;
    } else {
      double ix = x0[2 * j] - x0[2 * i], iy = x0[2 * j + 1] - x0[2 * i + 1];
      double id = sqrt(ix * ix + iy * iy);
      nx = ix / id; ny = iy / id;
    }
    dist_b[b] = d; nx_b[b] = nx; ny_b[b] = ny;
    double ext = d - rest[b];
    double tstrain = ext / rest[b];
    double bshear = 0.5 * (s[i] + s[j]);
    /* per-bond thresholds when given (strength scatter), else the config's */
    int sf = bshear > (eps_s_b ? eps_s_b[b] : prm[P_EPS_S]);
    int tf = tstrain > (eps_t_b ? eps_t_b[b] : prm[P_EPS_T]);
    if (alive[b] && (sf || tf)) {
      alive[b] = 0;
      fmode[b] = (uint8_t)(tf + 2 * sf);
      ftime[b] = prm[P_TIME];
      fext[b] = ext;
      fenergy[b] = (0.5 * kb) * (ext * ext);
      released += fenergy[b];
      double ov = two_r - d;
      if (ov < 0) ov = 0;
      acquired += (0.5 * kc) * (ov * ov);
      ++new_breaks;
    }
  }

  /* force-carrying pairs: bonds (alive spring or broken contact) + unbonded overlaps */
  int32_t *rowptr, *col;
  bond_csr(n, nb, bi, bj, &rowptr, &col);
  int nc = 0;
  pair_t *cp = grid_pairs(n, x, two_r, rowptr, col, &nc);
  pforce *pf = (pforce *)malloc(sizeof(pforce) * (size_t)(nb + nc > 0 ? nb + nc : 1));
  int npf = 0;
  double pair_diss = 0.0;
  /* pair order for the dissipation sum: merge bonds and contacts by (i, j) */
  int ib = 0, ic = 0;
  while (ib < nb || ic < nc) {
    int use_bond;
    if (ib >= nb) use_bond = 0;
    else if (ic >= nc) use_bond = 1;
    else use_bond = (bi[ib] < cp[ic].i) || (bi[ib] == cp[ic].i && bj[ib] < cp[ic].j);
    int i, j, is_alive;
    double d, nx, ny;
    if (use_bond) {
      i = bi[ib]; j = bj[ib];
      d = dist_b[ib]; nx = nx_b[ib]; ny = ny_b[ib];
      is_alive = alive[ib];
    } else {
      i = cp[ic].i; j = cp[ic].j;
      double dx = x[2 * j] - x[2 * i], dy = x[2 * j + 1] - x[2 * i + 1];
      d = sqrt(dx * dx + dy * dy);
      if (d > 0) {
        double dd = d > EPS64 ? d : EPS64;
        nx = dx / dd; ny = dy / dd;
      } else {
        double ix = x0[2 * j] - x0[2 * i], iy = x0[2 * j + 1] - x0[2 * i + 1];
        double id = sqrt(ix * ix + iy * iy);
        nx = ix / id; ny = iy / id;
      }
      is_alive = 0;
    }
    double mag;
    if (is_alive) {
      mag = kb * (d - rest[ib]);
      pf[npf].i = i; pf[npf].j = j; pf[npf].fx = mag * nx; pf[npf].fy = mag * ny; ++npf;
    } else {
      double ov = two_r - d;
      if (ov < 0) ov = 0;
      if (ov > 0) {
        double vn = (v[2 * j] - v[2 * i]) * nx + (v[2 * j + 1] - v[2 * i + 1]) * ny;
        double rep = kc * ov - cdamp * vn;
        if (rep < 0) rep = 0;
        mag = -rep;
        pf[npf].i = i; pf[npf].j = j; pf[npf].fx = mag * nx; pf[npf].fy = mag * ny; ++npf;
        pair_diss += (rep - kc * ov) * (-vn);
      }
    }
    if (use_bond) ++ib; else ++ic;
  }

  /* per-particle accumulation in index_add_ order */
  int32_t *deg = (int32_t *)calloc((size_t)n + 1, sizeof(int32_t));
  for (int t = 0; t < npf; ++t) { deg[pf[t].i + 1]++; deg[pf[t].j + 1]++; }
  for (int p = 0; p < n; ++p) deg[p + 1] += deg[p];
  contrib *cb = (contrib *)malloc(sizeof(contrib) * (size_t)(2 * npf > 0 ? 2 * npf : 1));
  int32_t *fill = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
  memcpy(fill, deg, sizeof(int32_t) * (size_t)n);
  for (int t = 0; t < npf; ++t) {
    contrib *a = &cb[fill[pf[t].i]++];
    a->partner = pf[t].j; a->group = 0; a->fx = pf[t].fx; a->fy = pf[t].fy;
    contrib *c2 = &cb[fill[pf[t].j]++];
    c2->partner = pf[t].i; c2->group = 1; c2->fx = -pf[t].fx; c2->fy = -pf[t].fy;
  }
  for (int p = 0; p < n; ++p) {
    qsort(cb + deg[p], (size_t)(deg[p + 1] - deg[p]), sizeof(contrib), cmp_contrib);
    double fx = (-drag) * v[2 * p], fy = (-drag) * v[2 * p + 1];
    for (int t = deg[p]; t < deg[p + 1]; ++t) { fx += cb[t].fx; fy += cb[t].fy; }
    force[2 * p] = fx;
    force[2 * p + 1] = fy;
  }

  /* tool (dem.py:294-324) and its dissipation (dem.py:327-341) */
  double rx = 0.0, ry = 0.0, tool_diss = 0.0, drag_sq = 0.0;
  const double tvx = prm[P_TOOL_VX], tvy = prm[P_TOOL_VY];
  for (int p = 0; p < n; ++p) {
    double tfx, tfy, cfx, cfy;
    if (hull == NULL) {
      double dx = x[2 * p] - prm[P_TOOL_X], dy = x[2 * p + 1] - prm[P_TOOL_Y];
      double d = sqrt(dx * dx + dy * dy);
      double nx, ny;
      if (d > 0) {
        double dd = d > EPS64 ? d : EPS64;
        nx = dx / dd; ny = dy / dd;
      } else {
        nx = 0.0; ny = -1.0;
      }
      double ov = r + prm[P_TOOL_R] - d;
      if (ov < 0) ov = 0;
      double sp = (v[2 * p] - tvx) * nx + (v[2 * p + 1] - tvy) * ny;
      double mag = kc * ov - cdamp * sp;
      if (mag < 0) mag = 0;
      if (!(ov > 0)) mag = 0;
      tfx = mag * nx; tfy = mag * ny;
      cfx = kc * ov * nx; cfy = kc * ov * ny;
    } else {
      double tv[2] = {tvx, tvy};
      hull_particle(x + 2 * p, v + 2 * p, hull, hull_m, tv, r, prm[P_TOOL_R], kc, cdamp, &tfx, &tfy);
      hull_particle(x + 2 * p, v + 2 * p, hull, hull_m, tv, r, prm[P_TOOL_R], kc, 0.0, &cfx, &cfy);
    }
    rx += tfx; ry += tfy;
    tool_diss += (tfx - cfx) * (v[2 * p] - tvx);
    tool_diss += (tfy - cfy) * (v[2 * p + 1] - tvy);
    drag_sq += v[2 * p] * v[2 * p];
    drag_sq += v[2 * p + 1] * v[2 * p + 1];
    force[2 * p] += tfx;
    force[2 * p + 1] += tfy;
    if (loads) {
      force[2 * p] += loads[2 * p];
      force[2 * p + 1] += loads[2 * p + 1];
    }
  }
  reaction[0] = -rx;
  reaction[1] = -ry;
  tool_diss = -tool_diss;
  scal[S_RELEASED] = released;
  scal[S_ACQUIRED] = acquired;
  scal[S_DAMP_POWER] = pair_diss + tool_diss + drag * drag_sq;
  scal[S_TOOL_POWER] = -(reaction[0] * tvx + reaction[1] * tvy);
  scal[S_NEW_BREAKS] = (double)new_breaks;

  free(s); free(dist_b); free(nx_b); free(ny_b); free(rowptr); free(col); free(cp);
  free(pf); free(deg); free(cb); free(fill);
}

void orc_plate_forces(const double *prm, int n, const double *x, const double *v,
                      const double *x0, int nb, const int32_t *bi, const int32_t *bj,
                      const double *rest, uint8_t *alive, uint8_t *fmode, double *ftime,
                      double *fext, double *fenergy, const double *loads, const double *hull,
                      int hull_m, double *force, double *reaction, double *scal)
{
  plate_forces(prm, n, x, v, x0, nb, bi, bj, rest, alive, fmode, ftime, fext, fenergy, loads,
               hull, hull_m, force, reaction, scal, NULL, NULL);
}

/* As orc_plate_forces with per-bond axial / shear thresholds (nb each). */
void orc_plate_forces_thr(const double *prm, int n, const double *x, const double *v,
                          const double *x0, int nb, const int32_t *bi, const int32_t *bj,
                          const double *rest, uint8_t *alive, uint8_t *fmode, double *ftime,
                          double *fext, double *fenergy, const double *loads,
                          const double *hull, int hull_m, double *force, double *reaction,
                          double *scal, const double *eps_t_b, const double *eps_s_b)
{
  plate_forces(prm, n, x, v, x0, nb, bi, bj, rest, alive, fmode, ftime, fext, fenergy, loads,
               hull, hull_m, force, reaction, scal, eps_t_b, eps_s_b);
}

/*
 * IceDEM.step (dem.py:347-382): forces, impulse ledger, semi-implicit Euler,
 * clamp of fixed particles, work ledger.  Returns 1 if the new state is
 * non-finite (the caller raises FloatingPointError), else 0.
 */
int orc_plate_step_thr(const double *prm, int n, double *x, double *v, const double *x0,
                       const uint8_t *fixed, int nb, const int32_t *bi, const int32_t *bj,
                       const double *rest, uint8_t *alive, uint8_t *fmode, double *ftime,
                       double *fext, double *fenergy, const double *loads, const double *hull,
                       int hull_m, double *acc, double *last, double *force_out,
                       const double *eps_t_b, const double *eps_s_b)
{
  const double dt = prm[P_DT], m = prm[P_MASS], drag = prm[P_DRAG];
  double *force = force_out;
  double reaction[2], scal[S_NSCAL];
  double *xold = (double *)malloc(sizeof(double) * (size_t)(2 * n > 0 ? 2 * n : 1));
  memcpy(xold, x, sizeof(double) * (size_t)(2 * n));
  plate_forces(prm, n, x, v, x0, nb, bi, bj, rest, alive, fmode, ftime, fext, fenergy, loads,
               hull, hull_m, force, reaction, scal, eps_t_b, eps_s_b);
  acc[A_RELEASE] += scal[S_RELEASED];
  acc[A_ACQUIRED] += scal[S_ACQUIRED];
  double sup[2] = {0.0, 0.0}, ext[2] = {0.0, 0.0}, vs[2] = {0.0, 0.0};
  for (int p = 0; p < n; ++p) {
    if (fixed[p]) { sup[0] += force[2 * p]; sup[1] += force[2 * p + 1]; }
    if (loads) { ext[0] += loads[2 * p]; ext[1] += loads[2 * p + 1]; }
    vs[0] += v[2 * p]; vs[1] += v[2 * p + 1];
  }
  sup[0] = -sup[0]; sup[1] = -sup[1];
  double dimp[2] = {(-drag) * vs[0], (-drag) * vs[1]};
  double timp[2] = {-reaction[0], -reaction[1]};
  acc[A_EXT_IMP_X] = acc[A_EXT_IMP_X] + dt * ext[0];
  acc[A_EXT_IMP_Y] = acc[A_EXT_IMP_Y] + dt * ext[1];
  acc[A_SUP_IMP_X] = acc[A_SUP_IMP_X] + dt * sup[0];
  acc[A_SUP_IMP_Y] = acc[A_SUP_IMP_Y] + dt * sup[1];
  acc[A_DRAG_IMP_X] = acc[A_DRAG_IMP_X] + dt * dimp[0];
  acc[A_DRAG_IMP_Y] = acc[A_DRAG_IMP_Y] + dt * dimp[1];
  acc[A_TOOL_IMP_X] = acc[A_TOOL_IMP_X] + dt * timp[0];
  acc[A_TOOL_IMP_Y] = acc[A_TOOL_IMP_Y] + dt * timp[1];
  double work = 0.0, moment = 0.0;
  int bad = 0;
  for (int p = 0; p < n; ++p) {
    for (int c = 0; c < 2; ++c) {
      double vn = v[2 * p + c] + (dt * force[2 * p + c]) / m;
      if (fixed[p]) vn = 0.0;
      v[2 * p + c] = vn;
      double xn = x[2 * p + c] + dt * vn;
      if (fixed[p]) xn = x0[2 * p + c];
      x[2 * p + c] = xn;
      if (!isfinite(xn) || !isfinite(vn)) bad = 1;
    }
    if (loads) {
      work += loads[2 * p] * (x[2 * p] - xold[2 * p]);
      work += loads[2 * p + 1] * (x[2 * p + 1] - xold[2 * p + 1]);
      moment += xold[2 * p] * loads[2 * p + 1] - xold[2 * p + 1] * loads[2 * p];
    }
  }
  acc[A_EXT_WORK] += work;
  acc[A_TOOL_WORK] += dt * scal[S_TOOL_POWER];
  acc[A_DAMPING] += dt * scal[S_DAMP_POWER];
  last[L_T0] = prm[P_TIME];
  last[L_T1] = prm[P_TIME] + dt;
  last[L_WORK] = work;
  last[L_SUP_X] = sup[0]; last[L_SUP_Y] = sup[1];
  last[L_REACT_X] = reaction[0]; last[L_REACT_Y] = reaction[1];
  last[L_EXT_X] = ext[0]; last[L_EXT_Y] = ext[1];
  last[L_MOMENT] = moment;
  free(xold);
  return bad;
}

int orc_plate_step(const double *prm, int n, double *x, double *v, const double *x0,
                   const uint8_t *fixed, int nb, const int32_t *bi, const int32_t *bj,
                   const double *rest, uint8_t *alive, uint8_t *fmode, double *ftime,
                   double *fext, double *fenergy, const double *loads, const double *hull,
                   int hull_m, double *acc, double *last, double *force_out)
{
  return orc_plate_step_thr(prm, n, x, v, x0, fixed, nb, bi, bj, rest, alive, fmode, ftime,
                            fext, fenergy, loads, hull, hull_m, acc, last, force_out, NULL, NULL);
}

/* mechanical_energy (dem.py:603-624) for the circular tool or the hull. */
void orc_mech_energy(const double *prm, int n, const double *x, const double *v, int nb,
                     const int32_t *bi, const int32_t *bj, const double *rest,
                     const uint8_t *alive, const double *hull, int hull_m, double *out5)
{
  const double r = prm[P_RADIUS], kb = prm[P_KB], kc = prm[P_KC];
  double bond = 0.0, contact = 0.0, tool = 0.0, kin = 0.0;
  int32_t *rowptr, *col;
  bond_csr(n, nb, bi, bj, &rowptr, &col);
  for (int b = 0; b < nb; ++b) {
    int i = bi[b], j = bj[b];
    double dx = x[2 * j] - x[2 * i], dy = x[2 * j + 1] - x[2 * i + 1];
    double d = sqrt(dx * dx + dy * dy);
    if (alive[b]) {
      double e = d - rest[b];
      bond += e * e;
    } else {
      double ov = 2 * r - d;
      if (ov > 0) contact += ov * ov;
    }
  }
  int nc = 0;
  pair_t *cp = grid_pairs(n, x, 2 * r, rowptr, col, &nc);
  for (int t = 0; t < nc; ++t) {
    int i = cp[t].i, j = cp[t].j;
    double dx = x[2 * j] - x[2 * i], dy = x[2 * j + 1] - x[2 * i + 1];
    double ov = 2 * r - sqrt(dx * dx + dy * dy);
    if (ov > 0) contact += ov * ov;
  }
  for (int p = 0; p < n; ++p) {
    double d;
    if (hull == NULL) {
      double dx = x[2 * p] - prm[P_TOOL_X], dy = x[2 * p + 1] - prm[P_TOOL_Y];
      d = sqrt(dx * dx + dy * dy);
    } else {
      d = INFINITY;
      for (int s = 0; s < hull_m - 1; ++s) {
        double sx = hull[2 * s], sy = hull[2 * s + 1];
        double gx = hull[2 * s + 2] - sx, gy = hull[2 * s + 3] - sy;
        double fr = ((x[2 * p] - sx) * gx + (x[2 * p + 1] - sy) * gy) / (gx * gx + gy * gy);
        if (fr < 0) fr = 0;
        if (fr > 1) fr = 1;
        double dx = x[2 * p] - (sx + fr * gx), dy = x[2 * p + 1] - (sy + fr * gy);
        double dd = sqrt(dx * dx + dy * dy);
        if (dd < d) d = dd;
      }
    }
    double ov = r + prm[P_TOOL_R] - d;
    if (ov > 0) tool += ov * ov;
    kin += v[2 * p] * v[2 * p] + v[2 * p + 1] * v[2 * p + 1];
  }
  out5[0] = 0.5 * prm[P_MASS] * kin;
  out5[1] = 0.5 * kb * bond;
  out5[2] = 0.5 * kc * contact;
  out5[3] = 0.5 * kc * tool;
  out5[4] = out5[0] + out5[1] + out5[2] + out5[3];
  free(rowptr); free(col); free(cp);
}

/* polyline_contact_forces (hull_contact.py:7-75) on its own: per-particle
 * force (n,2) from the nearest point of the polyline `vert` (m vertices, in
 * global position) moving with `hv`, and the hull reaction (2) = -sum of
 * the forces.  The plate step calls the same per-particle law. */
void orc_hull_forces(int n, const double *x, const double *v, const double *vert, int m,
                     const double *hv, double r_p, double r_h, double k, double c,
                     double *force, double *reaction)
{
  double sx = 0.0, sy = 0.0;
  for (int p = 0; p < n; ++p) {
    hull_particle(x + 2 * p, v + 2 * p, vert, m, hv, r_p, r_h, k, c, force + 2 * p,
                  force + 2 * p + 1);
    sx += force[2 * p];
    sy += force[2 * p + 1];
  }
  reaction[0] = -sx;
  reaction[1] = -sy;
}

/* Brute-force list of all pairs (i<j) closer than `cutoff`: the pair-set
 * oracle for the cell hash (O(n^2), small n only). */
int orc_pairs_within(int n, const double *x, double cutoff, int32_t *out_ij, int cap)
{
  int cnt = 0;
  for (int i = 0; i < n; ++i)
    for (int j = i + 1; j < n; ++j) {
      double dx = x[2 * j] - x[2 * i], dy = x[2 * j + 1] - x[2 * i + 1];
      if (sqrt(dx * dx + dy * dy) < cutoff) {
        if (cnt < cap) { out_ij[2 * cnt] = i; out_ij[2 * cnt + 1] = j; }
        ++cnt;
      }
    }
  return cnt;
}
