/*
 * FP64 CPU oracle for the unbonded Hertz-Mindlin disk bed (test infrastructure only).
 *
 * No single reference function computes a 2-D Hertz-Mindlin bed; this file
 * restates the 3-D floe-bed law of the reference in two dimensions:
 *   - normal: s_n = 2 E* sqrt(R* d), f_el = (2/3) s_n d, viscous term
 *     2 sqrt(5/6) beta sqrt(s_n m*) v_n with beta = -ln e / sqrt(ln^2 e + pi^2),
 *     clamped non-cohesive (src/tensordem/scale_ceiling/kernel.py:216-227,
 *     293-307);
 *   - tangential: Cundall-Strack spring xi += v_t dt, s_t = 8 G* sqrt(R* d),
 *     |s_t xi| <= mu f_n by scaling xi (kernel.py:308-322); the history of a
 *     pair is dropped when it separates (granular2d/engine.py:388-396,
 *     clear_ice/contact.py:316-323);
 *   - E* = E / (2 (1 - nu^2)), G* = G / (2 (2 - nu)), G = E / (2 (1 + nu))
 *     for equal materials (kernel.py:119-129); walls use the same E*, G*,
 *     R* = r, m* = m (kernel.py:362-399) and the scale cap/(s_t |xi|) of the
 *     pair law (the reference's wall branch divides by |xi| only);
 *   - floor y = y_floor and back wall x = x_left (granular2d/
 *     column_collapse.py:236-266: no gate wall once released), gravity;
 *   - a right wall (the floe bed's indenter, scale_ceiling/kernel.py:156-197,
 *     338-401, 514-519) at x = x_right + v_right * t during the step that
 *     starts at time t, inward normal -x, the law evaluated on the velocity
 *     relative to the wall; it doubles as the column's gate (v_right = 0);
 *   - semi-implicit Euler v += dt F / m, x += dt v as IceDEM.step
 *     (dem.py:363-365); the floe bed's velocity-Verlet
 *     (scale_ceiling/kernel.py:410-421) is driven from oracle/bed.py on this
 *     force evaluation (integrate = 0: springs committed, no update).
 * Per disk the force is gravity, then pair forces by ascending partner
 * index, then floor, then back wall.  (The reference's scatter order,
 * torch index_add_, is not specified; this order is the one the GPU kernel
 * follows so that FP64 runs compare bit for bit.)
 * Build with -ffp-contract=off.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

enum {
  B_ESTAR = 0, B_GSTAR, B_CDAMP, B_MU, B_MU_WALL, B_RHO_PI_T, B_GRAV_Y, B_DT, B_Y_FLOOR,
  B_X_LEFT, B_USE_FLOOR, B_USE_LEFT, B_X_RIGHT, B_V_RIGHT, B_USE_RIGHT, B_TIME, B_NPARAM
};
/* scalar outputs: powers absorbed (>= 0 when dissipating), counts, and the
 * power of the moving right wall on the disks */
enum {
  BS_DASH = 0, BS_FRIC, BS_WALL_DASH, BS_WALL_FRIC, BS_WALL_FX, BS_WALL_FY, BS_NPAIRS,
  BS_NWALL, BS_WALL_POWER, BS_NSCAL
};

typedef struct { int64_t key; int32_t idx; } bkeyed;

static int cmp_bkeyed(const void *a, const void *b)
{
  const bkeyed *u = (const bkeyed *)a, *v = (const bkeyed *)b;
  if (u->key != v->key) return u->key < v->key ? -1 : 1;
  return u->idx - v->idx;
}

typedef struct { int32_t i, j; } bpair;

static int cmp_bpair(const void *a, const void *b)
{
  const bpair *u = (const bpair *)a, *v = (const bpair *)b;
  if (u->i != v->i) return u->i - v->i;
  return u->j - v->j;
}

/* all pairs i<j with |x_j - x_i| < r_i + r_j (grid of cell 2 r_max) */
static bpair *bed_pairs(int n, const double *x, const double *r, int *count)
{
  double rmax = 0;
  for (int p = 0; p < n; ++p) if (r[p] > rmax) rmax = r[p];
  double h = 2 * rmax;
  bkeyed *k = (bkeyed *)malloc(sizeof(bkeyed) * (size_t)(n > 0 ? n : 1));
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
    k[m].key = cx[p] * 4294967296LL + cy[p];
    k[m].idx = p;
    ++m;
  }
  qsort(k, (size_t)m, sizeof(bkeyed), cmp_bkeyed);
  int cap = 64, cnt = 0;
  bpair *out = (bpair *)malloc(sizeof(bpair) * (size_t)cap);
  for (int p = 0; p < n; ++p) {
    if (cx[p] == INT64_MIN) continue;
    for (int ox = -1; ox <= 1; ++ox)
      for (int oy = -1; oy <= 1; ++oy) {
        int64_t key = (cx[p] + ox) * 4294967296LL + (cy[p] + oy);
        int lo = 0, hi = m;
        while (lo < hi) {
          int mid = (lo + hi) >> 1;
          if (k[mid].key < key) lo = mid + 1; else hi = mid;
        }
        for (int t = lo; t < m && k[t].key == key; ++t) {
          int q = k[t].idx;
          if (q <= p) continue;
          double dx = x[2 * q] - x[2 * p], dy = x[2 * q + 1] - x[2 * p + 1];
          double dist = sqrt(dx * dx + dy * dy);
          if (!((r[p] + r[q]) - dist > 0)) continue;
          if (cnt == cap) { cap *= 2; out = (bpair *)realloc(out, sizeof(bpair) * (size_t)cap); }
          out[cnt].i = p; out[cnt].j = q; ++cnt;
        }
      }
  }
  qsort(out, (size_t)cnt, sizeof(bpair), cmp_bpair);
  free(k); free(cx); free(cy);
  *count = cnt;
  return out;
}

typedef struct { int32_t partner; double fx, fy; } bcontrib;

static int cmp_bcontrib(const void *a, const void *b)
{
  const bcontrib *u = (const bcontrib *)a, *v = (const bcontrib *)b;
  return u->partner - v->partner;
}

static double mass_of(const double *prm, double r) { return prm[B_RHO_PI_T] * (r * r); }

/* wall contact: gap measured along the inward normal (nx, ny) */
static void wall(const double *prm, double r, double m, double gap, double nx, double ny,
                 const double *v, double *xi, int *active, double mu, double *fx, double *fy,
                 double *dash, double *fric)
{
  double d = r - gap;
  if (!(d > 0)) { xi[0] = 0.0; xi[1] = 0.0; *active = 0; *fx = 0; *fy = 0; return; }
  *active = 1;
  double sn = (2 * prm[B_ESTAR]) * sqrt(r * d);
  double fel = ((2.0 / 3.0) * sn) * d;
  double vn = v[0] * nx + v[1] * ny;
  double fn = fel - (prm[B_CDAMP] * sqrt(sn * m)) * vn;
  if (fn < 0) fn = 0;
  double st = (8 * prm[B_GSTAR]) * sqrt(r * d);
  double vtx = v[0] - vn * nx, vty = v[1] - vn * ny;
  double ex = xi[0] + vtx * prm[B_DT], ey = xi[1] + vty * prm[B_DT];
  double mag = sqrt(ex * ex + ey * ey);
  double cap = mu * fn;
  if (mag > 0) {
    double den = st * mag;
    double sc = cap / (den > 1e-300 ? den : 1e-300);
    if (sc > 1) sc = 1;
    ex = ex * sc; ey = ey * sc;
  }
  xi[0] = ex; xi[1] = ey;
  *fx = fn * nx - st * ex;
  *fy = fn * ny - st * ey;
  *dash += (fn - fel) * (-vn);
  *fric += st * (ex * v[0] + ey * v[1]);
}

/*
 * One bed step.  History in: hk (nh keys i*n+j, ascending), hx (nh,2).
 * History out: ok (cap_out), ox (cap_out,2); returns the new count, or -1 if
 * cap_out is too small.  xi_floor / xi_left (n,2) in/out.  force (n,2) out.
 * scal[BS_*] out (powers, wall reaction, counts).
 */
int orb_step(const double *prm, int n, double *x, double *v, const double *r, int nh,
             const int64_t *hk, const double *hx, int cap_out, int64_t *ok, double *ox,
             double *xi_floor, double *xi_left, double *xi_right, double *force, double *scal,
             int integrate)
{
  int np = 0;
  bpair *pr = bed_pairs(n, x, r, &np);
  if (np > cap_out) { free(pr); return -1; }
  double dash = 0, fric = 0, wdash = 0, wfric = 0, wfx = 0, wfy = 0, wpow = 0;
  int nwall = 0;
  bcontrib *cb = (bcontrib *)malloc(sizeof(bcontrib) * (size_t)(2 * np > 0 ? 2 * np : 1));
  int32_t *deg = (int32_t *)calloc((size_t)n + 1, sizeof(int32_t));
  double *pfx = (double *)malloc(sizeof(double) * (size_t)(np > 0 ? np : 1));
  double *pfy = (double *)malloc(sizeof(double) * (size_t)(np > 0 ? np : 1));
  int hp = 0;
  for (int t = 0; t < np; ++t) {
    const int i = pr[t].i, j = pr[t].j;
    const int64_t key = (int64_t)i * n + j;
    while (hp < nh && hk[hp] < key) ++hp;  /* sorted-key merge of the history */
    double e0x = 0, e0y = 0;
    if (hp < nh && hk[hp] == key) { e0x = hx[2 * hp]; e0y = hx[2 * hp + 1]; }
    double dx = x[2 * j] - x[2 * i], dy = x[2 * j + 1] - x[2 * i + 1];
    double dist = sqrt(dx * dx + dy * dy);
    double d = (r[i] + r[j]) - dist;
    double dd = dist > 1e-300 ? dist : 1e-300;
    double nx = dx / dd, ny = dy / dd;
    double reff = (r[i] * r[j]) / (r[i] + r[j]);
    double mi = mass_of(prm, r[i]), mj = mass_of(prm, r[j]);
    double meff = (mi * mj) / (mi + mj);
    double sn = (2 * prm[B_ESTAR]) * sqrt(reff * d);
    double fel = ((2.0 / 3.0) * sn) * d;
    double vrx = v[2 * i] - v[2 * j], vry = v[2 * i + 1] - v[2 * j + 1];
    double vn = vrx * nx + vry * ny;
    double fn = fel + (prm[B_CDAMP] * sqrt(sn * meff)) * vn;
    if (fn < 0) fn = 0;
    double st = (8 * prm[B_GSTAR]) * sqrt(reff * d);
    double vtx = vrx - vn * nx, vty = vry - vn * ny;
    double ex = e0x + vtx * prm[B_DT], ey = e0y + vty * prm[B_DT];
    double mag = sqrt(ex * ex + ey * ey);
    double cap = prm[B_MU] * fn;
    if (mag > 0) {
      double den = st * mag;
      double sc = cap / (den > 1e-300 ? den : 1e-300);
      if (sc > 1) sc = 1;
      ex = ex * sc; ey = ey * sc;
    }
    ok[t] = key;
    ox[2 * t] = ex; ox[2 * t + 1] = ey;
    pfx[t] = fn * nx + st * ex;  /* force on j */
    pfy[t] = fn * ny + st * ey;
    deg[i + 1]++; deg[j + 1]++;
    dash += (fn - fel) * vn;
    fric += st * (ex * vrx + ey * vry);
  }
  for (int p = 0; p < n; ++p) deg[p + 1] += deg[p];
  int32_t *fill = (int32_t *)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
  memcpy(fill, deg, sizeof(int32_t) * (size_t)n);
  for (int t = 0; t < np; ++t) {
    bcontrib *a = &cb[fill[pr[t].i]++];
    a->partner = pr[t].j; a->fx = -pfx[t]; a->fy = -pfy[t];
    bcontrib *b = &cb[fill[pr[t].j]++];
    b->partner = pr[t].i; b->fx = pfx[t]; b->fy = pfy[t];
  }
  int bad = 0;
  for (int p = 0; p < n; ++p) {
    qsort(cb + deg[p], (size_t)(deg[p + 1] - deg[p]), sizeof(bcontrib), cmp_bcontrib);
    const double m = mass_of(prm, r[p]);
    double fx = 0.0, fy = m * prm[B_GRAV_Y];
    for (int t = deg[p]; t < deg[p + 1]; ++t) { fx += cb[t].fx; fy += cb[t].fy; }
    double wx, wy;
    int act;
    if (prm[B_USE_FLOOR] != 0) {
      wall(prm, r[p], m, x[2 * p + 1] - prm[B_Y_FLOOR], 0.0, 1.0, v + 2 * p, xi_floor + 2 * p,
           &act, prm[B_MU_WALL], &wx, &wy, &wdash, &wfric);
      if (act) { fx += wx; fy += wy; wfx += wx; wfy += wy; ++nwall; }
    }
    if (prm[B_USE_LEFT] != 0) {
      wall(prm, r[p], m, x[2 * p] - prm[B_X_LEFT], 1.0, 0.0, v + 2 * p, xi_left + 2 * p, &act,
           prm[B_MU_WALL], &wx, &wy, &wdash, &wfric);
      if (act) { fx += wx; fy += wy; wfx += wx; wfy += wy; ++nwall; }
    }
    if (prm[B_USE_RIGHT] != 0) {
      const double xr = prm[B_X_RIGHT] + prm[B_V_RIGHT] * prm[B_TIME];
      const double vrel[2] = {v[2 * p] - prm[B_V_RIGHT], v[2 * p + 1]};
      wall(prm, r[p], m, xr - x[2 * p], -1.0, 0.0, vrel, xi_right + 2 * p, &act,
           prm[B_MU_WALL], &wx, &wy, &wdash, &wfric);
      if (act) {
        fx += wx; fy += wy; wfx += wx; wfy += wy; ++nwall;
        wpow += wx * prm[B_V_RIGHT];
      }
    }
    force[2 * p] = fx;
    force[2 * p + 1] = fy;
  }
  if (integrate) {
    const double dt = prm[B_DT];
    for (int p = 0; p < n; ++p) {
      const double m = mass_of(prm, r[p]);
      for (int c = 0; c < 2; ++c) {
        double vn = v[2 * p + c] + (dt * force[2 * p + c]) / m;
        v[2 * p + c] = vn;
        x[2 * p + c] = x[2 * p + c] + dt * vn;
        if (!isfinite(vn) || !isfinite(x[2 * p + c])) bad = 1;
      }
    }
  }
  scal[BS_DASH] = dash; scal[BS_FRIC] = fric; scal[BS_WALL_DASH] = wdash;
  scal[BS_WALL_FRIC] = wfric; scal[BS_WALL_FX] = wfx; scal[BS_WALL_FY] = wfy;
  scal[BS_NPAIRS] = (double)np; scal[BS_NWALL] = (double)nwall;
  scal[BS_WALL_POWER] = wpow;
  free(pr); free(cb); free(deg); free(pfx); free(pfy); free(fill);
  return bad ? -2 : np;
}

/* All pairs in contact (i<j), sorted: the contact-pair set at the given state. */
int orb_contacts(int n, const double *x, const double *r, int cap, int32_t *out_ij)
{
  int np = 0;
  bpair *pr = bed_pairs(n, x, r, &np);
  for (int t = 0; t < np && t < cap; ++t) { out_ij[2 * t] = pr[t].i; out_ij[2 * t + 1] = pr[t].j; }
  free(pr);
  return np;
}
