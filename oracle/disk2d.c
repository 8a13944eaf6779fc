/* FP64 CPU oracle for the 2-D rigid frictional disk assembly (test
 * infrastructure only: tests, the smoke check and the benchmark's CPU
 * baseline use it; the product never links it).
 *
 * Restates `DiskSystem._evaluate` and `DiskSystem.step` of
 * granular2d/engine.py:399-596 and 599-618 for one step:
 *
 *   contacts   pairs i < j with overlap r_i + r_j - |l| > 0, l = x_j - x_i
 *              under the minimum image on periodic axes (engine.py:298-350,
 *              411-418), in ascending key i*n + j; then wall segments and
 *              hull segments, disk-major, with the closest point
 *              t = clip(ap.ab / max(|ab|^2, 1e-30), 0, 1), overlap r - |d|,
 *              kept when overlap > 0 and |d| > 1e-12 (engine.py:352-375).
 *   state      spring xi (2) and rolling angle per contact, loaded by key
 *              (engine.py:388-397, 452-456): pairs by i*n + j, boundary
 *              contacts by n*n + k, k their rank among this step's wall then
 *              hull contacts (a boundary contact inherits the state of the
 *              previous step's boundary contact of the same rank).
 *   law        f_n = max(0, kn*ov - 2*zeta*sqrt(kn*m_eff)*v_n); xi += v_t dt,
 *              projected on the tangent; f_t = -k_t xi capped at mu f_n;
 *              rolling M_r = k_r theta_rel capped at mu_r r_eff f_n
 *              (engine.py:458-537).
 *   sums       force: pairs on i, pairs on j (negated), walls, hull, then
 *              gravity and drag; torque: lever crosses in the same order, then
 *              the rolling torques (engine.py:499-544) -- the order of
 *              numpy's np.add.at, so each disk's sums are bit-identical.
 *   step       semi-implicit Euler with the optional velocity decay and the
 *              periodic wrap x mod L (engine.py:599-618).
 *
 * Ledger and energy sums (ODS_*) add contacts in key order; numpy sums them
 * pairwise, so they agree with the reference to rounding, not bit for bit.
 * Compiled with -ffp-contract=off: every product and sum is rounded
 * separately, as numpy does.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

enum {
  OD_KN, OD_ZETA2, OD_KT, OD_MU, OD_MUR, OD_KR_COEF, OD_DT, OD_GX, OD_GY, OD_DRIFT_X,
  OD_DRIFT_Y, OD_DRAG_TIME, OD_DECAY, OD_LX, OD_LY, OD_USE_GRAVITY, OD_USE_DECAY, OD_NPARAM
};
enum {
  ODS_DASH, ODS_SLIDE, ODS_ROLL, ODS_EL_N, ODS_EL_T, ODS_EL_R, ODS_WALL_FX, ODS_WALL_FY,
  ODS_HULL_FX, ODS_HULL_FY, ODS_NPAIR, ODS_NWALL, ODS_NHULL, ODS_NSCAL
};
/* per-contact record fields (contact_forces, engine.py:546-575) */
enum {
  ODR_I, ODR_J, ODR_KIND, ODR_NX, ODR_NY, ODR_LX, ODR_LY, ODR_FX, ODR_FY, ODR_FN, ODR_FT,
  ODR_OV, ODR_PX, ODR_PY, ODR_NREC
};

typedef struct {
  int64_t i, j;   /* j = -1 for a boundary contact */
  int kind, seg;  /* 0 pair, 1 wall, 2 hull */
  double n[2], ov, pt[2], lever[2];
} Contact;

static double min_image(double d, double L) {
  /* d - L*round(d/L), numpy's round: half to even */
  return L > 0.0 ? d - L * nearbyint(d / L) : d;
}

static int cmp_i64(const void *a, const void *b) {
  int64_t x = *(const int64_t *)a, y = *(const int64_t *)b;
  return (x > y) - (x < y);
}

static int64_t find_key(const int64_t *keys, int64_t m, int64_t key) {
  int64_t lo = 0, hi = m;
  while (lo < hi) {
    int64_t mid = (lo + hi) / 2;
    if (keys[mid] < key) lo = mid + 1;
    else hi = mid;
  }
  return (lo < m && keys[lo] == key) ? lo : -1;
}

/* np.maximum(0.0, x): 0.0 unless x > 0 or x is NaN */
static double max0_first(double x) { return (0.0 >= x) ? 0.0 : x; }
/* np.maximum(x, 0.0) */
static double max0_second(double x) { return (x >= 0.0 || x != x) ? x : 0.0; }

/* Pair contacts in ascending key: a uniform grid of cells no smaller than
 * 2*r_max, so every overlapping pair lies in neighbouring cells. */
static int64_t pair_contacts(const double *p, int64_t n, const double *pos, const double *rad,
                             Contact **out) {
  const double Lx = p[OD_LX], Ly = p[OD_LY];
  double rmax = 0.0, lo[2] = {INFINITY, INFINITY}, hi[2] = {-INFINITY, -INFINITY};
  for (int64_t i = 0; i < n; ++i) {
    rmax = fmax(rmax, rad[i]);
    for (int a = 0; a < 2; ++a) {
      lo[a] = fmin(lo[a], pos[2 * i + a]);
      hi[a] = fmax(hi[a], pos[2 * i + a]);
    }
  }
  const double h0 = 2.0 * rmax;
  const double L[2] = {Lx, Ly};
  int64_t nc[2];
  double h[2];
  for (int a = 0; a < 2; ++a) {
    if (L[a] > 0.0) {
      nc[a] = (int64_t)floor(L[a] / h0);
      if (nc[a] < 1) nc[a] = 1;
      if (nc[a] > 4096) nc[a] = 4096;
      h[a] = L[a] / (double)nc[a];
    } else {
      nc[a] = (int64_t)floor((hi[a] - lo[a]) / h0) + 1;
      if (nc[a] > 4096) nc[a] = 4096;
      h[a] = fmax((hi[a] - lo[a]) / (double)nc[a], h0) * (1.0 + 1e-12);
      if (!(h[a] > 0.0)) h[a] = 1.0;
    }
  }
  const int64_t ncell = nc[0] * nc[1];
  int64_t *cell = malloc(sizeof(int64_t) * (size_t)n);
  int64_t *start = calloc((size_t)ncell + 1, sizeof(int64_t));
  int64_t *order = malloc(sizeof(int64_t) * (size_t)(n ? n : 1));
  for (int64_t i = 0; i < n; ++i) {
    int64_t c[2];
    for (int a = 0; a < 2; ++a) {
      double x = pos[2 * i + a];
      /* a periodic coordinate is wrapped into [0, L] first */
      double f = L[a] > 0.0 ? floor((x - L[a] * floor(x / L[a])) / h[a])
                            : floor((x - lo[a]) / h[a]);
      int64_t k = f >= 0.0 ? (f < (double)nc[a] ? (int64_t)f : nc[a] - 1) : 0;
      c[a] = k;
    }
    cell[i] = c[0] + nc[0] * c[1];
    start[cell[i] + 1]++;
  }
  for (int64_t c = 0; c < ncell; ++c) start[c + 1] += start[c];
  int64_t *fill = malloc(sizeof(int64_t) * (size_t)ncell);
  memcpy(fill, start, sizeof(int64_t) * (size_t)ncell);
  for (int64_t i = 0; i < n; ++i) order[fill[cell[i]]++] = i;
  free(fill);

  int64_t cap = 64 + 8 * n, m = 0;
  Contact *cs = malloc(sizeof(Contact) * (size_t)cap);
  int64_t *nb = malloc(sizeof(int64_t) * (size_t)(n ? n : 1));
  for (int64_t i = 0; i < n; ++i) {
    int64_t ci = cell[i], cx = ci % nc[0], cy = ci / nc[0];
    int64_t xs[3], ys[3];
    int nxs = 0, nys = 0;
    for (int d = -1; d <= 1; ++d) {
      int64_t k = cx + d;
      if (Lx > 0.0) { k %= nc[0]; if (k < 0) k += nc[0]; }
      else if (k < 0 || k >= nc[0]) continue;
      int dup = 0;
      for (int e = 0; e < nxs; ++e) dup |= xs[e] == k;
      if (!dup) xs[nxs++] = k;
    }
    for (int d = -1; d <= 1; ++d) {
      int64_t k = cy + d;
      if (Ly > 0.0) { k %= nc[1]; if (k < 0) k += nc[1]; }
      else if (k < 0 || k >= nc[1]) continue;
      int dup = 0;
      for (int e = 0; e < nys; ++e) dup |= ys[e] == k;
      if (!dup) ys[nys++] = k;
    }
    int64_t k = 0;
    for (int a = 0; a < nxs; ++a)
      for (int b = 0; b < nys; ++b) {
        int64_t c = xs[a] + nc[0] * ys[b];
        for (int64_t s = start[c]; s < start[c + 1]; ++s)
          if (order[s] > i) nb[k++] = order[s];
      }
    qsort(nb, (size_t)k, sizeof(int64_t), cmp_i64);
    for (int64_t e = 0; e < k; ++e) {
      int64_t j = nb[e];
      double l0 = min_image(pos[2 * j] - pos[2 * i], Lx);
      double l1 = min_image(pos[2 * j + 1] - pos[2 * i + 1], Ly);
      double dist = sqrt(l0 * l0 + l1 * l1);
      double dm = fmax(dist, 1e-30);
      double nij0 = l0 / dm, nij1 = l1 / dm;
      double ov = (rad[i] + rad[j]) - dist;
      if (!(ov > 0.0)) continue;
      if (m == cap) {
        cap *= 2;
        cs = realloc(cs, sizeof(Contact) * (size_t)cap);
      }
      Contact *c = &cs[m++];
      c->i = i;
      c->j = j;
      c->kind = 0;
      c->seg = -1;
      c->n[0] = -nij0;
      c->n[1] = -nij1;
      c->ov = ov;
      c->lever[0] = l0;
      c->lever[1] = l1;
      c->pt[0] = pos[2 * i] + rad[i] * nij0;
      c->pt[1] = pos[2 * i + 1] + rad[i] * nij1;
    }
  }
  free(nb);
  free(cell);
  free(start);
  free(order);
  *out = cs;
  return m;
}

/* Boundary contacts of one segment set, disk-major (np.nonzero order). */
static int64_t segment_contacts(int64_t n, const double *pos, const double *rad,
                                const double *seg, int ns, int kind, Contact **cs,
                                int64_t m, int64_t *cap) {
  for (int64_t d = 0; d < n; ++d)
    for (int s = 0; s < ns; ++s) {
      const double *g = seg + 4 * s;
      double ab0 = g[2] - g[0], ab1 = g[3] - g[1];
      double ap0 = pos[2 * d] - g[0], ap1 = pos[2 * d + 1] - g[1];
      double denom = fmax(ab0 * ab0 + ab1 * ab1, 1e-30);
      double t = (ap0 * ab0 + ap1 * ab1) / denom;
      if (t == t) t = fmin(fmax(t, 0.0), 1.0); /* np.clip keeps NaN */
      double cp0 = g[0] + t * ab0, cp1 = g[1] + t * ab1;
      double d0 = pos[2 * d] - cp0, d1 = pos[2 * d + 1] - cp1;
      double dist = sqrt(d0 * d0 + d1 * d1);
      double ov = rad[d] - dist;
      if (!(ov > 0.0 && dist > 1e-12)) continue;
      if (m == *cap) {
        *cap *= 2;
        *cs = realloc(*cs, sizeof(Contact) * (size_t)*cap);
      }
      Contact *c = &(*cs)[m++];
      c->i = d;
      c->j = -1;
      c->kind = kind;
      c->seg = s;
      c->n[0] = d0 / dist;
      c->n[1] = d1 / dist;
      c->ov = ov;
      c->pt[0] = cp0;
      c->pt[1] = cp1;
      c->lever[0] = cp0 - pos[2 * d];
      c->lever[1] = cp1 - pos[2 * d + 1];
    }
  return m;
}

/* One evaluation (and, with `advance`, one step) of the disk assembly.
 *
 * State in: `nst` contacts with sorted keys `st_key`, springs `st_xi` (2 per
 * contact) and rolling angles `st_roll`.  With `advance`, the new state is
 * written to `new_key/new_xi/new_roll` (capacity `cap`) and the disks are
 * integrated.  `rec` (optional, capacity `cap` rows of ODR_NREC) receives the
 * per-contact records.  Returns the number of contacts, -1 when `cap` is too
 * small.  `scal` receives ODS_* (ledger terms are this step's increments:
 * dashpot work already multiplied by dt, sliding and rolling losses >= 0). */
int64_t ord_eval(const double *p, int64_t n, double *pos, double *vel, double *theta,
                 double *omega, const double *rad, const double *mass, const double *inertia,
                 const double *walls, int n_walls, const double *wall_vel, const double *hull,
                 int n_hull, int64_t nst, const int64_t *st_key, const double *st_xi,
                 const double *st_roll, int64_t cap, int64_t *new_key, double *new_xi,
                 double *new_roll, double *rec, double *scal, int advance, int ledger) {
  Contact *cs = NULL;
  int64_t np_ = pair_contacts(p, n, pos, rad, &cs);
  int64_t ccap = np_ > 0 ? np_ : 1;
  Contact *all = malloc(sizeof(Contact) * (size_t)(ccap + 64));
  memcpy(all, cs, sizeof(Contact) * (size_t)np_);
  free(cs);
  ccap += 64;
  int64_t m = segment_contacts(n, pos, rad, walls, n_walls, 1, &all, np_, &ccap);
  int64_t nw = m - np_;
  m = segment_contacts(n, pos, rad, hull, n_hull, 2, &all, m, &ccap);
  int64_t nh = m - np_ - nw;
  for (int k = 0; k < ODS_NSCAL; ++k) scal[k] = 0.0;
  scal[ODS_NPAIR] = (double)np_;
  scal[ODS_NWALL] = (double)nw;
  scal[ODS_NHULL] = (double)nh;
  if (m > cap) {
    free(all);
    return -1;
  }
  const double kn = p[OD_KN], kt = p[OD_KT], dt = p[OD_DT];
  double *fx = calloc((size_t)(2 * n), sizeof(double));
  double *tq = calloc((size_t)n, sizeof(double));
  double *tot = malloc(sizeof(double) * (size_t)(2 * (m ? m : 1)));
  double *mr = malloc(sizeof(double) * (size_t)(m ? m : 1));
  double dash = 0.0, slide = 0.0, rollloss = 0.0, eln = 0.0, elt = 0.0, elr = 0.0;
  double wf[2] = {0.0, 0.0}, hf[2] = {0.0, 0.0};
  for (int64_t c = 0; c < m; ++c) {
    Contact *C = &all[c];
    const int64_t i = C->i;
    const int pair = C->kind == 0;
    const int64_t j = C->j;
    const int64_t key = pair ? i * n + j : n * n + (c - np_);
    const int64_t at = find_key(st_key, nst, key);
    double xi0 = 0.0, xi1 = 0.0, roll = 0.0;
    if (at >= 0) {
      xi0 = st_xi[2 * at];
      xi1 = st_xi[2 * at + 1];
      roll = st_roll[at];
    }
    double meff, reff;
    if (pair) {
      meff = (mass[i] * mass[j]) / (mass[i] + mass[j]);
      reff = rad[i] * rad[j] / (rad[i] + rad[j]);
    } else {
      meff = mass[i];
      reff = rad[i];
    }
    const double n0 = C->n[0], n1 = C->n[1];
    const double li0 = -rad[i] * n0, li1 = -rad[i] * n1;
    const double ro = pair ? rad[j] : 0.0;
    const double lj0 = ro * n0, lj1 = ro * n1;
    const double om_j = pair ? omega[j] : 0.0;
    /* v_i = vel_i + perp(omega_i * lever_i), perp(a) = (-a_y, a_x) */
    const double wi0 = omega[i] * li0, wi1 = omega[i] * li1;
    const double vi0 = vel[2 * i] + -wi1, vi1 = vel[2 * i + 1] + wi0;
    double vj0 = 0.0, vj1 = 0.0;
    if (pair) {
      const double wj0 = om_j * lj0, wj1 = om_j * lj1;
      vj0 = vel[2 * j] + -wj1;
      vj1 = vel[2 * j + 1] + wj0;
    } else if (C->kind == 1 && wall_vel) {
      vj0 = wall_vel[2 * C->seg];
      vj1 = wall_vel[2 * C->seg + 1];
    }
    const double vr0 = vi0 - vj0, vr1 = vi1 - vj1;
    const double vn = vr0 * n0 + vr1 * n1;
    const double cn = p[OD_ZETA2] * sqrt(kn * meff);
    const double fn = max0_first(kn * C->ov - cn * vn);
    const double vt0 = vr0 - vn * n0, vt1 = vr1 - vn * n1;
    xi0 = xi0 + vt0 * dt;
    xi1 = xi1 + vt1 * dt;
    const double xn = xi0 * n0 + xi1 * n1;
    xi0 = xi0 - xn * n0;
    xi1 = xi1 - xn * n1;
    double ft0 = -kt * xi0, ft1 = -kt * xi1;
    double ftm = sqrt(ft0 * ft0 + ft1 * ft1);
    const double cap_t = p[OD_MU] * fn;
    const double scale = ftm > cap_t ? cap_t / fmax(ftm, 1e-30) : 1.0;
    slide += cap_t * max0_second(ftm - cap_t) / kt;
    ft0 *= scale;
    ft1 *= scale;
    xi0 *= scale;
    xi1 *= scale;
    ftm *= scale;
    const double t0 = fn * n0 + ft0, t1 = fn * n1 + ft1;
    tot[2 * c] = t0;
    tot[2 * c + 1] = t1;
    mr[c] = 0.0;
    if (p[OD_MUR] > 0.0) {
      roll = roll + (omega[i] - om_j) * dt;
      const double kr = p[OD_KR_COEF] * (reff * reff);
      double m_r = kr * roll;
      const double cap_r = p[OD_MUR] * reff * fn;
      const double am = fabs(m_r);
      const double sr = am > cap_r ? cap_r / fmax(am, 1e-30) : 1.0;
      rollloss += cap_r * max0_second(am - cap_r) / kr;
      m_r *= sr;
      roll *= sr;
      mr[c] = m_r;
      elr += kr * (roll * roll);
    }
    dash += (fn - kn * C->ov) * vn;
    eln += C->ov * C->ov;
    elt += xi0 * xi0 + xi1 * xi1;
    if (C->kind == 1) {
      wf[0] += t0;
      wf[1] += t1;
    } else if (C->kind == 2) {
      hf[0] += t0;
      hf[1] += t1;
    }
    if (advance) {
      new_key[c] = key;
      new_xi[2 * c] = xi0;
      new_xi[2 * c + 1] = xi1;
      new_roll[c] = roll;
    }
    if (rec) {
      double *r = rec + (size_t)c * ODR_NREC;
      r[ODR_I] = (double)i;
      r[ODR_J] = (double)j;
      r[ODR_KIND] = (double)C->kind;
      r[ODR_NX] = n0;
      r[ODR_NY] = n1;
      r[ODR_LX] = C->lever[0];
      r[ODR_LY] = C->lever[1];
      r[ODR_FX] = t0;
      r[ODR_FY] = t1;
      r[ODR_FN] = fn;
      r[ODR_FT] = ftm;
      r[ODR_OV] = C->ov;
      r[ODR_PX] = C->pt[0];
      r[ODR_PY] = C->pt[1];
    }
  }
  /* np.add.at order: pairs on i, pairs on j, walls, hull */
  for (int64_t c = 0; c < np_; ++c) {
    const int64_t i = all[c].i;
    fx[2 * i] += tot[2 * c];
    fx[2 * i + 1] += tot[2 * c + 1];
  }
  for (int64_t c = 0; c < np_; ++c) {
    const int64_t j = all[c].j;
    fx[2 * j] += -tot[2 * c];
    fx[2 * j + 1] += -tot[2 * c + 1];
  }
  for (int64_t c = 0; c < np_; ++c) {
    const int64_t i = all[c].i;
    const double li0 = -rad[i] * all[c].n[0], li1 = -rad[i] * all[c].n[1];
    tq[i] += li0 * tot[2 * c + 1] - li1 * tot[2 * c];
  }
  for (int64_t c = 0; c < np_; ++c) {
    const int64_t j = all[c].j;
    const double lj0 = rad[j] * all[c].n[0], lj1 = rad[j] * all[c].n[1];
    tq[j] += lj0 * -tot[2 * c + 1] - lj1 * -tot[2 * c];
  }
  for (int64_t c = np_; c < m; ++c) {
    const int64_t i = all[c].i;
    fx[2 * i] += tot[2 * c];
    fx[2 * i + 1] += tot[2 * c + 1];
    const double li0 = -rad[i] * all[c].n[0], li1 = -rad[i] * all[c].n[1];
    tq[i] += li0 * tot[2 * c + 1] - li1 * tot[2 * c];
  }
  if (p[OD_MUR] > 0.0) {
    for (int64_t c = 0; c < np_; ++c) tq[all[c].i] += -mr[c];
    for (int64_t c = 0; c < np_; ++c) tq[all[c].j] += mr[c];
    for (int64_t c = np_; c < m; ++c) tq[all[c].i] += -mr[c];
  }
  for (int64_t d = 0; d < n; ++d) {
    if (p[OD_USE_GRAVITY] != 0.0) {
      fx[2 * d] = fx[2 * d] + mass[d] * p[OD_GX];
      fx[2 * d + 1] = fx[2 * d + 1] + mass[d] * p[OD_GY];
    }
    if (p[OD_DRAG_TIME] > 0.0) {
      fx[2 * d] = fx[2 * d] + mass[d] * (p[OD_DRIFT_X] - vel[2 * d]) / p[OD_DRAG_TIME];
      fx[2 * d + 1] = fx[2 * d + 1] + mass[d] * (p[OD_DRIFT_Y] - vel[2 * d + 1]) / p[OD_DRAG_TIME];
    }
  }
  scal[ODS_DASH] = ledger ? dash * dt : 0.0;
  scal[ODS_SLIDE] = ledger ? slide : 0.0;
  scal[ODS_ROLL] = ledger ? rollloss : 0.0;
  scal[ODS_EL_N] = 0.5 * kn * eln;
  scal[ODS_EL_T] = 0.5 * kt * elt;
  scal[ODS_EL_R] = p[OD_MUR] > 0.0 ? 0.5 * elr : 0.0;
  scal[ODS_WALL_FX] = -wf[0];
  scal[ODS_WALL_FY] = -wf[1];
  scal[ODS_HULL_FX] = -hf[0];
  scal[ODS_HULL_FY] = -hf[1];
  if (advance) {
    for (int64_t d = 0; d < n; ++d) {
      vel[2 * d] += dt * (fx[2 * d] / mass[d]);
      vel[2 * d + 1] += dt * (fx[2 * d + 1] / mass[d]);
      omega[d] += dt * (tq[d] / inertia[d]);
      if (p[OD_USE_DECAY] != 0.0) {
        vel[2 * d] *= p[OD_DECAY];
        vel[2 * d + 1] *= p[OD_DECAY];
        omega[d] *= p[OD_DECAY];
      }
      pos[2 * d] += dt * vel[2 * d];
      pos[2 * d + 1] += dt * vel[2 * d + 1];
      theta[d] += dt * omega[d];
      for (int a = 0; a < 2; ++a) {
        const double L = a ? p[OD_LY] : p[OD_LX];
        if (L > 0.0) {
          /* numpy remainder: fmod, then shift a negative remainder by L */
          double r = fmod(pos[2 * d + a], L);
          if (r != 0.0) {
            if ((r < 0.0) != (L < 0.0)) r += L;
          } else {
            r = copysign(0.0, L);
          }
          pos[2 * d + a] = r;
        }
      }
    }
  }
  free(fx);
  free(tq);
  free(tot);
  free(mr);
  free(all);
  return m;
}
