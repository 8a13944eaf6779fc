/*
 * FP64 CPU oracle for the 3-D bonded-sphere solver IceDEM3D (test
 * infrastructure only; never linked into the product).
 *
 * Restates one force evaluation of src/tensordem/dem3d.py for the options the
 * device solver supports:
 *   - bonds: axial spring k_b (L - L0) along the current unit vector, the
 *     rest vector's unit when the spheres coincide (dem3d.py:1327-1350);
 *   - fracture before forces: tensile strain (L - L0)/L0 above the scalar or
 *     per-bond Weibull threshold, or the mean of the two end spheres'
 *     objective shear strain above the shear threshold; failure mode,
 *     time, extension and energy 0.5 k_b ext^2 recorded (dem3d.py:984-1009);
 *   - objective shear strain of a sphere: F = (sum q r^T)(sum r r^T)^-1 over
 *     every bond of the sphere, broken or not, E = (F^T F - I)/2,
 *     sqrt(2/3 dev E : dev E), zero when sum r r^T is rank deficient
 *     (dem3d.py:875-901);
 *   - contacts: every pair closer than 2r whose bond is absent or broken,
 *     linear k_c or Hertz secant k = (4/3) E* sqrt(R* d) normal spring, the
 *     dashpot gamma_n = 2 zeta sqrt(k m_eff) (zeta from the restitution
 *     coefficient, m_eff = m/2, or m when exactly one sphere is fixed) or
 *     the absolute dashpot c, clamped non-cohesive (dem3d.py:1352-1447);
 *   - Cundall-Strack friction: xi += v_t dt, F_t = k_t xi scaled onto
 *     mu F_n, stored back as F_t / k_t, k_t constant or the Mindlin
 *     8 G* sqrt(R* d) (dem3d.py:1420-1493); a pair that separates (or whose
 *     bond is alive) starts again from xi = 0;
 *   - the spherical tool (linear or rigid Hertz spring, absolute dashpot on
 *     the relative normal velocity, dem3d.py:1529-1543), the horizontal top
 *     and bottom platens and the four lateral platens (dem3d.py:1545-1637),
 *     drag, gravity and nodal loads;
 *   - semi-implicit Euler with fixed spheres clamped (dem3d.py:1644-1660).
 * Hertz forces against rigid surfaces are c d sqrt(d) (the reference writes
 * c d^1.5; both are within an ulp).
 *
 * Contact partners come from a sorted cell list of width 2r (every pair closer
 * than 2r lies in adjacent cells); I3_BRUTE tests every pair instead.
 *
 * Per sphere the force is summed as: drag, gravity, load, forward bonds
 * (partner above, ascending), backward bonds (partner below, ascending),
 * contacts by ascending partner (friction, then normal), tool, top and
 * bottom platen, lateral platens.  The reference sums with torch index_add_
 * in the same bond order but visits contacts in its cell-dictionary order;
 * this fixed order is the one the device kernel follows, so FP64 runs
 * compare bit for bit.  Build with -ffp-contract=off.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

enum {
  D3_RADIUS = 0, D3_MASS, D3_DT, D3_DRAG, D3_KB, D3_KC, D3_KT, D3_MU, D3_EPS_T, D3_EPS_S,
  D3_HZ_PAIR, D3_HZ_PAIR_U, D3_HZ_RIGID, D3_HZ_KT, D3_ZETA2, D3_GAMMA_MOBILE, D3_GAMMA_PINNED, D3_CDAMP,
  D3_TOOL_R, D3_TOOL_X0, D3_TOOL_V = D3_TOOL_X0 + 3, D3_PLATEN_K = D3_TOOL_V + 3, D3_PLATEN_C,
  D3_TOP_Z0, D3_BOT_Z0, D3_TOP_V, D3_BOT_V, D3_LAT_K, D3_LAT_C, D3_MG, D3_TIME,
  D3_LAT_POS, D3_LAT_VEL = D3_LAT_POS + 4, D3_NPARAM = D3_LAT_VEL + 4
};
enum {
  I3_HERTZ = 0, I3_DAMP, I3_FRICTION, I3_MINDLIN_KT, I3_TOP, I3_BOT, I3_LATERAL, I3_GRAVITY,
  I3_UPDATE_FRACTURE, I3_UPDATE_HISTORY, I3_OP, I3_KH, I3_BRUTE, I3_NIPARAM
};
enum { DAMP_NONE = 0, DAMP_RESTITUTION = 1, DAMP_ABSOLUTE = 2 };
enum { OP_STEP = 0, OP_FORCES = 1, OP_ENERGY = 2 };
/* scalar outputs of one evaluation */
enum {
  O3_REACTION = 0, O3_TOP_Z = 3, O3_BOT_Z, O3_LAT, O3_FRICTION_POWER = O3_LAT + 4, O3_CONTACTS,
  O3_BREAKS, O3_STATUS, O3_NOUT
};
#define ST_NONFINITE 1
#define ST_HIST_OVERFLOW 16

static const double kEps = 2.220446049250313e-16; /* torch.finfo(float64).eps */

/* inverse of the symmetric reference covariance of every sphere, from its
 * cofactors; rank deficient (flag 1) when det <= 1e-10 trace^3 */
void orc3_covinv(int n, int nb, const int32_t *pairs, const double *x0, double *inv,
                 uint8_t *degenerate)
{
  double *cov = (double *)calloc((size_t)n * 9 + 1, sizeof(double));
  /* dem3d.py:883-885: index_add over i (bond order), then over j */
  for (int pass = 0; pass < 2; ++pass)
    for (int b = 0; b < nb; ++b) {
      int i = pairs[2 * b], j = pairs[2 * b + 1];
      int p = pass == 0 ? i : j;
      double r[3];
      for (int k = 0; k < 3; ++k) r[k] = x0[3 * j + k] - x0[3 * i + k];
      for (int a = 0; a < 3; ++a)
        for (int c = 0; c < 3; ++c) cov[9 * p + 3 * a + c] += r[a] * r[c];
    }
  for (int p = 0; p < n; ++p) {
    const double *m = cov + 9 * p;
    double a = m[0], b = m[1], c = m[2], d = m[4], e = m[5], f = m[8];
    double A = d * f - e * e, B = c * e - b * f, C = b * e - c * d;
    double D = a * f - c * c, E = b * c - a * e, F = a * d - b * b;
    double det = a * A + b * B + c * C;
    double tr = a + d + f;
    double *o = inv + 9 * p;
    if (!(det > 1e-10 * tr * tr * tr)) {
      degenerate[p] = 1;
      for (int k = 0; k < 9; ++k) o[k] = 0.0;
      continue;
    }
    degenerate[p] = 0;
    o[0] = A / det; o[1] = B / det; o[2] = C / det;
    o[3] = B / det; o[4] = D / det; o[5] = E / det;
    o[6] = C / det; o[7] = E / det; o[8] = F / det;
  }
  free(cov);
}

typedef struct {
  int *fwd_start, *fwd;  /* bonds with i == p, ascending j */
  int *bwd_start, *bwd;  /* bonds with j == p, ascending i */
} adjacency;

static void build_adjacency(int n, int nb, const int32_t *pairs, adjacency *g)
{
  g->fwd_start = (int *)calloc((size_t)n + 1, sizeof(int));
  g->bwd_start = (int *)calloc((size_t)n + 1, sizeof(int));
  g->fwd = (int *)malloc(sizeof(int) * (size_t)(nb > 0 ? nb : 1));
  g->bwd = (int *)malloc(sizeof(int) * (size_t)(nb > 0 ? nb : 1));
  for (int b = 0; b < nb; ++b) {
    g->fwd_start[pairs[2 * b] + 1]++;
    g->bwd_start[pairs[2 * b + 1] + 1]++;
  }
  for (int p = 0; p < n; ++p) {
    g->fwd_start[p + 1] += g->fwd_start[p];
    g->bwd_start[p + 1] += g->bwd_start[p];
  }
  int *fo = (int *)calloc((size_t)n + 1, sizeof(int)), *bo = (int *)calloc((size_t)n + 1, sizeof(int));
  /* bonds are sorted by (i, j): bonds of one i come in ascending j, and a
   * sphere's backward bonds come in ascending i */
  for (int b = 0; b < nb; ++b) {
    int i = pairs[2 * b], j = pairs[2 * b + 1];
    g->fwd[g->fwd_start[i] + fo[i]++] = b;
    g->bwd[g->bwd_start[j] + bo[j]++] = b;
  }
  free(fo);
  free(bo);
}

static void free_adjacency(adjacency *g)
{
  free(g->fwd_start); free(g->bwd_start); free(g->fwd); free(g->bwd);
}

static double norm3(const double *d) { return sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]); }

/* objective shear strain of every sphere (dem3d.py:875-901) */
void orc3_shear(int n, int nb, const int32_t *pairs, const double *x, const double *x0,
                const double *inv, const uint8_t *degenerate, double *shear)
{
  adjacency g;
  build_adjacency(n, nb, pairs, &g);
  for (int p = 0; p < n; ++p) {
    double cr[9] = {0};
    for (int pass = 0; pass < 2; ++pass) {
      int s0 = pass == 0 ? g.fwd_start[p] : g.bwd_start[p];
      int s1 = pass == 0 ? g.fwd_start[p + 1] : g.bwd_start[p + 1];
      for (int s = s0; s < s1; ++s) {
        int b = pass == 0 ? g.fwd[s] : g.bwd[s];
        int i = pairs[2 * b], j = pairs[2 * b + 1];
        double q[3], r[3];
        for (int k = 0; k < 3; ++k) {
          q[k] = x[3 * j + k] - x[3 * i + k];
          r[k] = x0[3 * j + k] - x0[3 * i + k];
        }
        for (int a = 0; a < 3; ++a)
          for (int c = 0; c < 3; ++c) cr[3 * a + c] += q[a] * r[c];
      }
    }
    if (degenerate[p]) {
      shear[p] = 0.0;
      continue;
    }
    const double *m = inv + 9 * p;
    double F[9], e[9];
    for (int a = 0; a < 3; ++a)
      for (int c = 0; c < 3; ++c)
        F[3 * a + c] = cr[3 * a] * m[c] + cr[3 * a + 1] * m[3 + c] + cr[3 * a + 2] * m[6 + c];
    for (int a = 0; a < 3; ++a)
      for (int c = 0; c < 3; ++c) {
        double ftf = F[a] * F[c] + F[3 + a] * F[3 + c] + F[6 + a] * F[6 + c];
        e[3 * a + c] = 0.5 * (ftf - (a == c ? 1.0 : 0.0));
      }
    double mean = (e[0] + e[4] + e[8]) / 3.0;
    double s = 0.0;
    for (int a = 0; a < 3; ++a)
      for (int c = 0; c < 3; ++c) {
        double d = e[3 * a + c] - (a == c ? mean : 0.0);
        s += d * d;
      }
    shear[p] = sqrt((2.0 / 3.0) * s);
  }
  free_adjacency(&g);
}

typedef struct {
  double *xi;    /* n * kh * 3 */
  int32_t *q;    /* n * kh */
  int32_t *cnt;  /* n */
} history;

static const double *hist_find(const history *h, int kh, int p, int q)
{
  for (int s = 0; s < h->cnt[p]; ++s)
    if (h->q[(size_t)p * kh + s] == q) return h->xi + ((size_t)p * kh + s) * 3;
  return NULL;
}

typedef struct { int64_t c[3]; int32_t idx; } cellkey;

static int cmp_cellkey(const void *a, const void *b)
{
  const cellkey *u = (const cellkey *)a, *v = (const cellkey *)b;
  for (int k = 0; k < 3; ++k)
    if (u->c[k] != v->c[k]) return u->c[k] < v->c[k] ? -1 : 1;
  return u->idx - v->idx;
}

static int cmp_int(const void *a, const void *b)
{
  int u = *(const int *)a, v = *(const int *)b;
  return (u > v) - (u < v);
}

static int64_t cell_of(double x, double w)
{
  double c = floor(x / w);
  if (c < -1e15) c = -1e15;
  if (c > 1e15) c = 1e15;
  return (int64_t)c;
}

/* first entry of the sorted keys with cell >= c */
static int lower_cell(const cellkey *k, int m, const int64_t *c)
{
  int lo = 0, hi = m;
  while (lo < hi) {
    int mid = (lo + hi) / 2;
    cellkey probe = {{c[0], c[1], c[2]}, -1};
    if (cmp_cellkey(&k[mid], &probe) < 0) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

/* the candidate partners of p: every sphere of the 27 cells of width 2r
 * around it (brute: every sphere), ascending */
static int candidates(int n, const double *x, double w, const cellkey *keys, int m, int p,
                      int brute, int *out)
{
  int c = 0;
  if (brute) {
    for (int q = 0; q < n; ++q)
      if (q != p) out[c++] = q;
    return c;
  }
  if (!isfinite(x[3 * p]) || !isfinite(x[3 * p + 1]) || !isfinite(x[3 * p + 2])) return 0;
  int64_t cp[3] = {cell_of(x[3 * p], w), cell_of(x[3 * p + 1], w), cell_of(x[3 * p + 2], w)};
  for (int dz = -1; dz <= 1; ++dz)
    for (int dy = -1; dy <= 1; ++dy)
      for (int dx = -1; dx <= 1; ++dx) {
        int64_t cc[3] = {cp[0] + dx, cp[1] + dy, cp[2] + dz};
        for (int e = lower_cell(keys, m, cc); e < m; ++e) {
          const cellkey *k = keys + e;
          if (k->c[0] != cc[0] || k->c[1] != cc[1] || k->c[2] != cc[2]) break;
          if (k->idx != p) out[c++] = k->idx;
        }
      }
  qsort(out, (size_t)c, sizeof(int), cmp_int);
  return c;
}

/*
 * One evaluation.  x, v: state (advanced in place by OP_STEP); alive and the
 * failure arrays: bond state (committed when I3_UPDATE_FRACTURE); hist_in /
 * hist_out: friction springs per sphere, canonical orientation (lower index
 * first), sorted by partner; force (n,3) written by OP_FORCES and OP_STEP;
 * pair_energy (n) by OP_ENERGY (each pair to its lower sphere).
 * Returns the status bits.
 */
int orc3_eval(int n, int nb, const double *dp, const int *ip, double *x, double *v,
              const double *x0, const uint8_t *fixed, const int32_t *pairs, uint8_t *alive,
              const double *thr, const double *inv, const uint8_t *degenerate,
              uint8_t *fmode, double *ftime, double *fext, double *fenergy, const double *loads,
              const double *hxi_in, const int32_t *hq_in, const int32_t *hcnt_in,
              double *hxi_out, int32_t *hq_out, int32_t *hcnt_out, double *force,
              double *pair_energy, double *out)
{
  const double r = dp[D3_RADIUS], m = dp[D3_MASS], dt = dp[D3_DT], time = dp[D3_TIME];
  const int op = ip[I3_OP], kh = ip[I3_KH];
  const int commit = ip[I3_UPDATE_FRACTURE] && op != OP_ENERGY;
  const int upd_hist = ip[I3_UPDATE_HISTORY] && op != OP_ENERGY && ip[I3_FRICTION];
  int status = 0;
  for (int k = 0; k < O3_NOUT; ++k) out[k] = 0.0;

  adjacency g;
  build_adjacency(n, nb, pairs, &g);
  double *shear = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
  double *ext = (double *)malloc(sizeof(double) * (size_t)(nb > 0 ? nb : 1));
  double *bf = (double *)malloc(sizeof(double) * 3 * (size_t)(nb > 0 ? nb : 1));
  orc3_shear(n, nb, pairs, x, x0, inv, degenerate, shear);

  /* bond extensions, fracture commit, bond forces (dem3d.py:1327-1350, 984-1009) */
  for (int b = 0; b < nb; ++b) {
    int i = pairs[2 * b], j = pairs[2 * b + 1];
    double d[3], d0[3];
    for (int k = 0; k < 3; ++k) {
      d[k] = x[3 * j + k] - x[3 * i + k];
      d0[k] = x0[3 * j + k] - x0[3 * i + k];
    }
    double dist = norm3(d), l0 = norm3(d0);
    double e = dist - l0;
    ext[b] = e;
    if (commit && alive[b]) {
      double t = thr ? thr[b] : dp[D3_EPS_T];
      int tensile = e / l0 > t;
      int sh = 0.5 * (shear[i] + shear[j]) > dp[D3_EPS_S];
      if (tensile || sh) {
        alive[b] = 0;
        fmode[b] = (uint8_t)(tensile + 2 * sh);
        ftime[b] = time;
        fext[b] = e;
        fenergy[b] = 0.5 * dp[D3_KB] * (e * e);
        out[O3_BREAKS] += 1.0;
      }
    }
    double mag = dp[D3_KB] * e;
    for (int k = 0; k < 3; ++k) {
      double nk = dist > kEps ? d[k] / dist : d0[k] / (l0 > kEps ? l0 : kEps);
      bf[3 * b + k] = mag * nk;
    }
  }

  history hin = {(double *)hxi_in, (int32_t *)hq_in, (int32_t *)hcnt_in};
  history hout = {hxi_out, hq_out, hcnt_out};
  if (upd_hist)
    for (int p = 0; p < n; ++p) hout.cnt[p] = 0;

  const double two_r = 2.0 * r;
  const int brute = ip[I3_BRUTE];
  cellkey *keys = (cellkey *)malloc(sizeof(cellkey) * (size_t)(n > 0 ? n : 1));
  int *cand = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
  int nkeys = 0;
  if (!brute) {
    for (int p = 0; p < n; ++p) {
      if (!isfinite(x[3 * p]) || !isfinite(x[3 * p + 1]) || !isfinite(x[3 * p + 2])) continue;
      cellkey k = {{cell_of(x[3 * p], two_r), cell_of(x[3 * p + 1], two_r),
                    cell_of(x[3 * p + 2], two_r)}, p};
      keys[nkeys++] = k;
    }
    qsort(keys, (size_t)nkeys, sizeof(cellkey), cmp_cellkey);
  }
  for (int p = 0; p < n; ++p) {
    double f[3];
    for (int k = 0; k < 3; ++k) f[k] = -dp[D3_DRAG] * v[3 * p + k];
    if (ip[I3_GRAVITY]) f[2] += -dp[D3_MG];
    if (loads)
      for (int k = 0; k < 3; ++k) f[k] += loads[3 * p + k];
    for (int s = g.fwd_start[p]; s < g.fwd_start[p + 1]; ++s) {
      int b = g.fwd[s];
      if (alive[b])
        for (int k = 0; k < 3; ++k) f[k] += bf[3 * b + k];
    }
    for (int s = g.bwd_start[p]; s < g.bwd_start[p + 1]; ++s) {
      int b = g.bwd[s];
      if (alive[b])
        for (int k = 0; k < 3; ++k) f[k] -= bf[3 * b + k];
    }
    double epair = 0.0;
    const int nc = candidates(n, x, two_r, keys, nkeys, p, brute, cand);
    for (int e = 0; e < nc; ++e) {
      const int q = cand[e];
      int i = p < q ? p : q, j = p < q ? q : p;
      double cd[3];
      for (int k = 0; k < 3; ++k) cd[k] = x[3 * j + k] - x[3 * i + k];
      if (fabs(cd[0]) >= two_r || fabs(cd[1]) >= two_r || fabs(cd[2]) >= two_r) continue;
      double cdist = norm3(cd);
      double ov = two_r - cdist;
      if (!(ov > 0.0)) continue;
      /* alive bond between i and j? */
      int bonded = 0;
      for (int s = g.fwd_start[i]; s < g.fwd_start[i + 1]; ++s)
        if (pairs[2 * g.fwd[s] + 1] == j) bonded = alive[g.fwd[s]];
      if (bonded) continue;
      if (op == OP_ENERGY) {
        /* (0.5 k_c) d^2, or the Hertz potential (8/15) E* sqrt(R*) d^2 sqrt(d)
         * (dem3d.py:1765-1775) */
        if (p == i) {
          if (ip[I3_HERTZ]) {
            epair += dp[D3_HZ_PAIR_U] * (ov * ov * sqrt(ov));
          } else {
            epair += (0.5 * dp[D3_KC]) * (ov * ov);
          }
        }
        continue;
      }
      double cn[3], rv[3];
      double den = cdist > kEps ? cdist : kEps;
      for (int k = 0; k < 3; ++k) {
        cn[k] = cd[k] / den;
        rv[k] = v[3 * j + k] - v[3 * i + k];
      }
      double vn = rv[0] * cn[0] + rv[1] * cn[1] + rv[2] * cn[2];
      int pinned = fixed[i] ^ fixed[j];
      double meff = pinned ? m : m / 2.0;
      double kloc = ip[I3_HERTZ] ? dp[D3_HZ_PAIR] * sqrt(ov) : dp[D3_KC];
      double gam = 0.0;
      if (ip[I3_DAMP] == DAMP_ABSOLUTE) gam = dp[D3_CDAMP];
      else if (ip[I3_DAMP] == DAMP_RESTITUTION)
        gam = ip[I3_HERTZ] ? dp[D3_ZETA2] * sqrt(kloc * meff)
                           : (pinned ? dp[D3_GAMMA_PINNED] : dp[D3_GAMMA_MOBILE]);
      double mag = kloc * ov - gam * vn;
      if (mag < 0.0) mag = 0.0;
      double ft[3] = {0.0, 0.0, 0.0};
      if (ip[I3_FRICTION]) {
        const double *old = hist_find(&hin, kh, p, q);
        double vt[3], td[3];
        for (int k = 0; k < 3; ++k) {
          vt[k] = rv[k] - vn * cn[k];
          td[k] = old ? old[k] : 0.0;
        }
        if (ip[I3_UPDATE_HISTORY])
          for (int k = 0; k < 3; ++k) td[k] = td[k] + vt[k] * dt;
        double kt = ip[I3_MINDLIN_KT] ? dp[D3_HZ_KT] * sqrt(ov) : dp[D3_KT];
        double raw[3];
        for (int k = 0; k < 3; ++k) raw[k] = kt * td[k];
        double fm = norm3(raw);
        double lim = dp[D3_MU] * mag;
        double scale = fm > lim ? lim / (fm > 1e-15 ? fm : 1e-15) : 1.0;
        for (int k = 0; k < 3; ++k) ft[k] = raw[k] * scale;
        if (p == i)
          out[O3_FRICTION_POWER] += ft[0] * (v[3 * i] - v[3 * j]) + ft[1] * (v[3 * i + 1] - v[3 * j + 1])
                                    + ft[2] * (v[3 * i + 2] - v[3 * j + 2]);
        if (upd_hist) {
          int c = hout.cnt[p];
          if (c >= kh) {
            status |= ST_HIST_OVERFLOW;
          } else {
            hout.q[(size_t)p * kh + c] = q;
            for (int k = 0; k < 3; ++k) hout.xi[((size_t)p * kh + c) * 3 + k] = ft[k] / kt;
            hout.cnt[p] = c + 1;
          }
        }
      }
      if (p == i) out[O3_CONTACTS] += 1.0;
      double cf[3];
      for (int k = 0; k < 3; ++k) cf[k] = -mag * cn[k];
      if (p == i) {
        for (int k = 0; k < 3; ++k) f[k] += ft[k];
        for (int k = 0; k < 3; ++k) f[k] += cf[k];
      } else {
        for (int k = 0; k < 3; ++k) f[k] -= ft[k];
        for (int k = 0; k < 3; ++k) f[k] -= cf[k];
      }
    }
    if (op == OP_ENERGY) {
      pair_energy[p] = epair;
      continue;
    }
    /* tool (dem3d.py:1529-1543) */
    {
      double td[3], tn[3];
      for (int k = 0; k < 3; ++k)
        td[k] = x[3 * p + k] - (dp[D3_TOOL_X0 + k] + time * dp[D3_TOOL_V + k]);
      double tdist = norm3(td);
      double den = tdist > kEps ? tdist : kEps;
      for (int k = 0; k < 3; ++k) tn[k] = td[k] / den;
      if (!(tdist > kEps)) { tn[0] = 0.0; tn[1] = 0.0; tn[2] = -1.0; }
      double tov = r + dp[D3_TOOL_R] - tdist;
      if (tov < 0.0) tov = 0.0;
      double tvn = (v[3 * p] - dp[D3_TOOL_V]) * tn[0] + (v[3 * p + 1] - dp[D3_TOOL_V + 1]) * tn[1]
                   + (v[3 * p + 2] - dp[D3_TOOL_V + 2]) * tn[2];
      double el = ip[I3_HERTZ] ? dp[D3_HZ_RIGID] * (tov * sqrt(tov)) : dp[D3_KC] * tov;
      double tm = el - dp[D3_CDAMP] * tvn;
      if (tm < 0.0) tm = 0.0;
      if (!(tov > 0.0)) tm = 0.0;
      for (int k = 0; k < 3; ++k) {
        double tf = tm * tn[k];
        out[O3_REACTION + k] -= tf;
        f[k] += tf;
      }
    }
    /* top and bottom platens (dem3d.py:1545-1593) */
    if (ip[I3_TOP]) {
      double z = dp[D3_TOP_Z0] + time * dp[D3_TOP_V];
      double ov = x[3 * p + 2] + r - z;
      if (ov < 0.0) ov = 0.0;
      double el = ip[I3_HERTZ] ? dp[D3_HZ_RIGID] * (ov * sqrt(ov)) : dp[D3_PLATEN_K] * ov;
      double mg = el + dp[D3_PLATEN_C] * (v[3 * p + 2] - dp[D3_TOP_V]);
      if (mg < 0.0) mg = 0.0;
      if (!(ov > 0.0)) mg = 0.0;
      f[2] += -mg;
      out[O3_TOP_Z] += mg;
    }
    if (ip[I3_BOT]) {
      double z = dp[D3_BOT_Z0] + time * dp[D3_BOT_V];
      double ov = z + r - x[3 * p + 2];
      if (ov < 0.0) ov = 0.0;
      double el = ip[I3_HERTZ] ? dp[D3_HZ_RIGID] * (ov * sqrt(ov)) : dp[D3_PLATEN_K] * ov;
      double mg = el - dp[D3_PLATEN_C] * (v[3 * p + 2] - dp[D3_BOT_V]);
      if (mg < 0.0) mg = 0.0;
      if (!(ov > 0.0)) mg = 0.0;
      f[2] += mg;
      out[O3_BOT_Z] -= mg;
    }
    /* lateral platens x_min, x_max, y_min, y_max (dem3d.py:1595-1637) */
    if (ip[I3_LATERAL]) {
      const double *lp = dp + D3_LAT_POS, *lv = dp + D3_LAT_VEL;
      for (int w = 0; w < 4; ++w) {
        int ax = w < 2 ? 0 : 1;
        double xc = x[3 * p + ax], vc = v[3 * p + ax] - lv[ax == 0 ? 0 : 2];
        double ov = (w & 1) ? xc + r - lp[w] : lp[w] - xc + r;
        if (ov < 0.0) ov = 0.0;
        double mg = (w & 1) ? dp[D3_LAT_K] * ov + dp[D3_LAT_C] * vc
                            : dp[D3_LAT_K] * ov - dp[D3_LAT_C] * vc;
        if (mg < 0.0) mg = 0.0;
        if (!(ov > 0.0)) mg = 0.0;
        if (w & 1) { f[ax] -= mg; out[O3_LAT + w] += mg; }
        else { f[ax] += mg; out[O3_LAT + w] -= mg; }
      }
    }
    for (int k = 0; k < 3; ++k) force[3 * p + k] = f[k];
  }
  if (op == OP_STEP) {
    /* dem3d.py:1647-1653 */
    for (int p = 0; p < n; ++p)
      for (int k = 0; k < 3; ++k) {
        double acc = force[3 * p + k] / m;
        if (fixed[p]) acc = 0.0;
        double vv = v[3 * p + k] + acc * dt;
        if (fixed[p]) vv = 0.0;
        double xx = x[3 * p + k] + vv * dt;
        if (fixed[p]) xx = x0[3 * p + k];
        v[3 * p + k] = vv;
        x[3 * p + k] = xx;
        if (!isfinite(xx) || !isfinite(vv)) status |= ST_NONFINITE;
      }
  }
  free(shear); free(ext); free(bf); free(keys); free(cand);
  free_adjacency(&g);
  out[O3_STATUS] = status;
  return status;
}
