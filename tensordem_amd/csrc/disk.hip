// Rigid frictional disks with rotation: the step of the reference's
// DiskSystem (src/tensordem/granular2d/engine.py:399-596) on the device.
//
// Two kernels per step, one thread per disk:
//
//   disk_prep_kernel   counts each disk's wall and hull contacts at the
//                      current positions; the last block to finish turns the
//                      per-block counts into exclusive offsets (the rank of a
//                      boundary contact among all boundary contacts, which is
//                      how the reference keys their spring state,
//                      engine.py:449-453).  When a candidate search is due, it
//                      also drops every disk into the cell table.
//   disk_step_kernel   (search step: builds the disk's sorted candidate list
//                      from the 3x3 cells, carrying each surviving partner's
//                      spring over), evaluates every contact of the disk in
//                      the reference's order, integrates, and writes the new
//                      spring state; the last block reduces the per-block
//                      ledger partials in a fixed order.
//
// Every pair is evaluated by both of its disks from the same operands in the
// same order (i = lower index), so both copies of its spring state stay
// identical and no disk writes another disk's data: no float atomics, and
// each disk's force and torque are summed exactly as numpy's np.add.at sums
// them (pairs where it is i by ascending j, pairs where it is j by ascending
// i, walls, hull, then gravity and drag; the rolling torques after all lever
// torques), so FP64 trajectories match the reference bit for bit.
#include "common.h"

namespace tdm {
namespace {

constexpr int kDiskCap = 24;  // candidate slots per disk held in registers
constexpr int kDch = 13;      // ledger partial channels
enum {
  D_DASH, D_SLIDE, D_ROLL, D_OV2, D_XI2, D_KROLL2, D_WFX, D_WFY, D_HFX, D_HFY, D_NPAIR,
  D_NWALL, D_NHULL
};

// device-side step state (in the workspace)
struct DiskDyn {
  int launch;    // step-kernel launches so far (parity of the rebuild flags)
  int adv;       // steps advanced (parity of the boundary spring state)
  int lp;        // list buffer holding the current candidate lists
  int cellp;     // cell table the next search bins into
  int rebuilds;
  int status;
  int nst;       // boundary contacts in the stored spring state
  int nw, nh;    // wall / hull contacts at the current positions
  int pad0;
  unsigned done1, done3;  // finished blocks of the current prep / step launch
  int flag[2];   // candidate search due at the launch of that parity
  int pad1[2];
};

struct DiskWs {
  DiskDyn *dyn;
  int *cand;          // [2][S][N]
  double2 *hxi;       // [2][S][N] spring of each candidate pair
  double *hroll;      // [2][S][N] rolling angle of each candidate pair
  int *cnt;           // [2][N] candidates per disk
  double2 *xb;        // [N] position at the last search
  int *ccount;        // [2][ncells]
  int *centry;        // [2][ncells][cell_cap]
  double4 *wst;       // [2][wcap] boundary springs by rank: xi, roll
  int *blkcnt;        // [nblocks] packed wall | hull << 16 contacts per block
  int2 *blkoff;       // [nblocks] exclusive wall, hull offsets
  double *part;       // [nblocks][kDch]
  double *mr;         // [S + n_walls + n_hull][N] rolling torque terms
};

void carve(char *base, int n, int ncells, int cap, int S, int nw, int nh, DiskWs *w,
           size_t *total)
{
  const int nb = (n + kThreads - 1) / kThreads;
  const size_t wcap = (size_t)n * (size_t)(nw + nh) + 1;
  size_t off = 0;
  auto take = [&](size_t bytes) {
    char *p = base ? base + off : reinterpret_cast<char *>(off);
    off += align256(bytes);
    return p;
  };
  w->dyn = reinterpret_cast<DiskDyn *>(take(sizeof(DiskDyn)));
  w->cand = reinterpret_cast<int *>(take(sizeof(int) * 2 * (size_t)S * n));
  w->hxi = reinterpret_cast<double2 *>(take(sizeof(double2) * 2 * (size_t)S * n));
  w->hroll = reinterpret_cast<double *>(take(sizeof(double) * 2 * (size_t)S * n));
  w->cnt = reinterpret_cast<int *>(take(sizeof(int) * 2 * (size_t)n));
  w->xb = reinterpret_cast<double2 *>(take(sizeof(double2) * (size_t)n));
  w->ccount = reinterpret_cast<int *>(take(sizeof(int) * 2 * (size_t)ncells));
  w->centry = reinterpret_cast<int *>(take(sizeof(int) * 2 * (size_t)ncells * cap));
  w->wst = reinterpret_cast<double4 *>(take(sizeof(double4) * 2 * wcap));
  w->blkcnt = reinterpret_cast<int *>(take(sizeof(int) * (size_t)nb));
  w->blkoff = reinterpret_cast<int2 *>(take(sizeof(int2) * (size_t)nb));
  w->part = reinterpret_cast<double *>(take(sizeof(double) * kDch * (size_t)nb));
  w->mr = reinterpret_cast<double *>(take(sizeof(double) * (size_t)(S + nw + nh) * n));
  *total = off;
}

struct DiskArgs {
  const double2 *x, *v;
  const double *om;
  double2 *xo, *vo;
  double *omo;
  double *theta;
  const double *radius, *mass, *inertia;
  const double *walls, *hull, *wvel;
  double *out, *rec;
  DiskWs w;
  int ledger;
};

__device__ inline double min_image(double d, double L) { return L > 0.0 ? d - L * rint(d / L) : d; }

// np.maximum(0.0, x) and np.maximum(x, 0.0): NaN propagates, ties keep 0.0
__device__ inline double max0_first(double x) { return (0.0 >= x) ? 0.0 : x; }
__device__ inline double max0_second(double x) { return (x >= 0.0 || x != x) ? x : 0.0; }

// Closest point of a segment (engine.py:359-380): contact when the overlap
// r - |d| is positive and |d| > 1e-12; n points from the segment to the disk.
struct SegHit {
  double n0, n1, ov, cp0, cp1;
};

__device__ inline bool seg_hit(const double *g, double px, double py, double r, SegHit &h)
{
  const double a0 = g[0], a1 = g[1];
  const double ab0 = g[2] - a0, ab1 = g[3] - a1;
  const double ap0 = px - a0, ap1 = py - a1;
  const double denom = fmax(ab0 * ab0 + ab1 * ab1, 1e-30);
  double t = (ap0 * ab0 + ap1 * ab1) / denom;
  if (t == t) t = fmin(fmax(t, 0.0), 1.0);
  const double cp0 = a0 + t * ab0, cp1 = a1 + t * ab1;
  const double d0 = px - cp0, d1 = py - cp1;
  const double dist = sqrt(d0 * d0 + d1 * d1);
  const double ov = r - dist;
  if (!(ov > 0.0 && dist > 1e-12)) return false;
  h.n0 = d0 / dist;
  h.n1 = d1 / dist;
  h.ov = ov;
  h.cp0 = cp0;
  h.cp1 = cp1;
  return true;
}

__device__ inline int boundary_counts(const tdm_disk_params &P, const DiskArgs &A, double px,
                                      double py, double r)
{
  int cw = 0, chh = 0;
  SegHit h;
  for (int s = 0; s < P.n_walls; ++s) cw += seg_hit(A.walls + 4 * s, px, py, r, h);
  for (int s = 0; s < P.n_hull; ++s) chh += seg_hit(A.hull + 4 * s, px, py, r, h);
  return cw | (chh << 16);
}

__device__ inline int cell_coord(const tdm_disk_params &P, int a, double x)
{
  const int nc = P.ncell[a];
  if (P.period[a] > 0.0) {
    // wrap into [0, L] first, so a rounding of x/h never skips a cell
    const double L = P.period[a];
    const double f = floor((x - L * floor(x / L)) / P.cell_h[a]);
    if (!(f >= 0.0)) return 0;
    return f >= (double)(nc - 1) ? nc - 1 : (int)f;
  }
  double f = floor((x - P.grid_origin[a]) / P.cell_h[a]);
  if (!(f >= 0.0)) return 0;
  return f >= (double)(nc - 1) ? nc - 1 : (int)f;
}

// Neighbouring cells along one axis: wrapped and de-duplicated on a periodic
// axis (fewer than three cells there), clipped otherwise.
__device__ inline int axis_cells(const tdm_disk_params &P, int a, int c, int (&out)[3])
{
  const int nc = P.ncell[a];
  int k = 0;
  for (int d = -1; d <= 1; ++d) {
    int q = c + d;
    if (P.period[a] > 0.0) {
      q = ((q % nc) + nc) % nc;
    } else if (q < 0 || q >= nc) {
      continue;
    }
    bool dup = false;
    for (int e = 0; e < k; ++e) dup |= out[e] == q;
    if (!dup) out[k++] = q;
  }
  return k;
}

__device__ inline double atomic_load_f64(const double *p)
{
  unsigned long long u = __hip_atomic_load(reinterpret_cast<const unsigned long long *>(p),
                                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return __longlong_as_double((long long)u);
}

__device__ inline int atomic_load_i32(const int *p)
{
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// One contact of the law (engine.py:455-534).  (n0, n1) points from j to
// i, (vr0, vr1) = v_i - v_j at the contact, omrel = omega_i - omega_j.
// Updates the spring (xi, roll) and returns the force on i in (t0, t1) and
// the rolling torque m_r (on i: -m_r).
struct LawOut {
  double t0, t1, m_r, fn, ft, slide, rloss, kr;
};

__device__ inline LawOut contact_law(const tdm_disk_params &P, double n0, double n1, double ov,
                                     double meff, double reff, double vr0, double vr1,
                                     double omrel, double &xi0, double &xi1, double &roll)
{
  LawOut o;
  const double vn = vr0 * n0 + vr1 * n1;
  const double cn = P.zeta2 * sqrt(P.kn * meff);
  const double fn = max0_first(P.kn * ov - cn * vn);
  const double vt0 = vr0 - vn * n0, vt1 = vr1 - vn * n1;
  xi0 = xi0 + vt0 * P.dt;
  xi1 = xi1 + vt1 * P.dt;
  const double xn = xi0 * n0 + xi1 * n1;
  xi0 = xi0 - xn * n0;
  xi1 = xi1 - xn * n1;
  double ft0 = -P.kt * xi0, ft1 = -P.kt * xi1;
  const double ftm = sqrt(ft0 * ft0 + ft1 * ft1);
  const double cap = P.mu * fn;
  const double sc = ftm > cap ? cap / fmax(ftm, 1e-30) : 1.0;
  o.slide = cap * max0_second(ftm - cap) / P.kt;
  ft0 *= sc;
  ft1 *= sc;
  xi0 *= sc;
  xi1 *= sc;
  o.ft = ftm * sc;
  o.fn = fn;
  o.t0 = fn * n0 + ft0;
  o.t1 = fn * n1 + ft1;
  o.m_r = 0.0;
  o.rloss = 0.0;
  o.kr = 0.0;
  if (P.mu_rolling > 0.0) {
    roll = roll + omrel * P.dt;
    const double kr = P.kr_coef * (reff * reff);
    double m_r = kr * roll;
    const double cap_r = P.mu_rolling * reff * fn;
    const double am = fabs(m_r);
    const double sr = am > cap_r ? cap_r / fmax(am, 1e-30) : 1.0;
    o.rloss = cap_r * max0_second(am - cap_r) / kr;
    m_r *= sr;
    roll *= sr;
    o.m_r = m_r;
    o.kr = kr;
  }
  return o;
}

__device__ inline void put_record(double *rec, size_t at, double i, double j, double kind,
                                  double n0, double n1, double l0, double l1, const LawOut &o,
                                  double ov, double p0, double p1)
{
  double *r = rec + at * TDM_DISK_NREC;
  r[0] = i; r[1] = j; r[2] = kind; r[3] = n0; r[4] = n1; r[5] = l0; r[6] = l1;
  r[7] = o.t0; r[8] = o.t1; r[9] = o.fn; r[10] = o.ft; r[11] = ov; r[12] = p0; r[13] = p1;
}

__device__ inline void clear_record(double *rec, size_t at)
{
  rec[at * TDM_DISK_NREC + 2] = -1.0;
}

__global__ __launch_bounds__(kThreads) void disk_prep_kernel(tdm_disk_params P, DiskArgs A)
{
  __shared__ int bsum;
  __shared__ int last;
  DiskDyn *dyn = A.w.dyn;
  const int p = blockIdx.x * kThreads + threadIdx.x;
  const int launch = dyn->launch;
  const bool rebuild = dyn->flag[launch & 1] != 0;
  const int cellp = dyn->cellp;
  if (threadIdx.x == 0) bsum = 0;
  __syncthreads();
  if (p < P.n) {
    const double2 xp = A.x[p];
    const int packed = boundary_counts(P, A, xp.x, xp.y, A.radius[p]);
    if (packed) atomicAdd(&bsum, packed);
    if (rebuild) {
      const int ncells = P.ncell[0] * P.ncell[1];
      const int c = cell_coord(P, 0, xp.x) + P.ncell[0] * cell_coord(P, 1, xp.y);
      const size_t cc = (size_t)cellp * ncells + c;
      const int slot = atomicAdd(&A.w.ccount[cc], 1);
      if (slot < P.cell_cap) A.w.centry[cc * P.cell_cap + slot] = p;
      else atomicOr(&dyn->status, TDM_ST_CELL_OVERFLOW);
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    A.w.blkcnt[blockIdx.x] = bsum;
    __threadfence();
    last = atomicAdd(&dyn->done1, 1u) == gridDim.x - 1;
  }
  __syncthreads();
  if (!last) return;
  __threadfence();
  // the last block: exclusive offsets of the wall and hull counts per block
  const int nb = gridDim.x;
  const int per = (nb + kThreads - 1) / kThreads;
  const int b0 = threadIdx.x * per, b1 = min(nb, b0 + per);
  int sw = 0, sh = 0;
  for (int b = b0; b < b1; ++b) {
    const int v = atomic_load_i32(&A.w.blkcnt[b]);
    sw += v & 0xFFFF;
    sh += v >> 16;
  }
  // inclusive scan of (sw, sh) over the threads, 2 x 256 ints in LDS
  __shared__ int sw_s[kThreads], sh_s[kThreads];
  sw_s[threadIdx.x] = sw;
  sh_s[threadIdx.x] = sh;
  __syncthreads();
  for (int d = 1; d < kThreads; d <<= 1) {
    const int aw = threadIdx.x >= d ? sw_s[threadIdx.x - d] : 0;
    const int ah = threadIdx.x >= d ? sh_s[threadIdx.x - d] : 0;
    __syncthreads();
    sw_s[threadIdx.x] += aw;
    sh_s[threadIdx.x] += ah;
    __syncthreads();
  }
  int ow = sw_s[threadIdx.x] - sw, oh = sh_s[threadIdx.x] - sh;
  for (int b = b0; b < b1; ++b) {
    const int v = atomic_load_i32(&A.w.blkcnt[b]);
    A.w.blkoff[b] = make_int2(ow, oh);
    ow += v & 0xFFFF;
    oh += v >> 16;
  }
  if (threadIdx.x == kThreads - 1) {
    dyn->nw = sw_s[kThreads - 1];
    dyn->nh = sh_s[kThreads - 1];
    dyn->done1 = 0u;
  }
}

template <bool kAdvance>
__global__ __launch_bounds__(kThreads) void disk_step_kernel(tdm_disk_params P, DiskArgs A)
{
  __shared__ double red[4][kCh];
  __shared__ int scan[kThreads];
  __shared__ int last;
  DiskDyn *dyn = A.w.dyn;
  const DiskWs &W = A.w;
  const int N = P.n, S = P.cand_slots;
  const int p = blockIdx.x * kThreads + threadIdx.x;
  const bool live = p < N;
  const int launch = dyn->launch;
  const int par = launch & 1;
  const bool rebuild = dyn->flag[par] != 0;
  const int lp = dyn->lp;
  const int lq = rebuild ? lp ^ 1 : lp;
  const int cellp = dyn->cellp;
  const int adv = dyn->adv;
  const int nst = dyn->nst, nw_tot = dyn->nw;
  const int ncells = P.ncell[0] * P.ncell[1];
  double ch[kDch];
#pragma unroll
  for (int k = 0; k < kDch; ++k) ch[k] = 0.0;

  if (rebuild) {  // clear the idle cell table for the next search
    int *cc = W.ccount + (size_t)(cellp ^ 1) * ncells;
    for (int c = blockIdx.x * kThreads + threadIdx.x; c < ncells; c += gridDim.x * kThreads)
      cc[c] = 0;
  }

  double2 xp = make_double2(0.0, 0.0), vp = xp;
  double omp = 0.0, rp = 0.0, mp = 1.0;
  int packed = 0;
  if (live) {
    xp = A.x[p];
    vp = A.v[p];
    omp = A.om[p];
    rp = A.radius[p];
    mp = A.mass[p];
    packed = boundary_counts(P, A, xp.x, xp.y, rp);
  }
  // exclusive prefix of the packed boundary counts inside the block
  scan[threadIdx.x] = packed;
  __syncthreads();
  for (int d = 1; d < kThreads; d <<= 1) {
    const int a = threadIdx.x >= d ? scan[threadIdx.x - d] : 0;
    __syncthreads();
    scan[threadIdx.x] += a;
    __syncthreads();
  }
  const int excl = scan[threadIdx.x] - packed;
  const int2 boff = W.blkoff[blockIdx.x];
  const int rank_w0 = boff.x + (excl & 0xFFFF);
  const int rank_h0 = nw_tot + boff.y + (excl >> 16);

  // ---- candidate search (engine.py:298-357 gives the pair set) -------------
  if (live && rebuild) {
    int nl[kDiskCap];
    int k = 0;
    const double reach2 = P.reach * P.reach;
    const int cx = cell_coord(P, 0, xp.x), cy = cell_coord(P, 1, xp.y);
    int xs[3], ys[3];
    const int nxs = axis_cells(P, 0, cx, xs), nys = axis_cells(P, 1, cy, ys);
    for (int a = 0; a < nxs; ++a)
      for (int b = 0; b < nys; ++b) {
        const size_t c = (size_t)cellp * ncells + xs[a] + (size_t)P.ncell[0] * ys[b];
        const int m = min(W.ccount[c], P.cell_cap);
        for (int s = 0; s < m; ++s) {
          const int q = W.centry[c * P.cell_cap + s];
          if (q == p) continue;
          const double2 xq = A.x[q];
          const double l0 = min_image(xq.x - xp.x, P.period[0]);
          const double l1 = min_image(xq.y - xp.y, P.period[1]);
          if (!(l0 * l0 + l1 * l1 < reach2)) continue;
          if (k == S) {
            atomicOr(&dyn->status, TDM_ST_CAND_OVERFLOW);
            continue;
          }
          int e = k++;
          for (; e > 0 && nl[e - 1] > q; --e) nl[e] = nl[e - 1];
          nl[e] = q;
        }
      }
    // carry the springs of partners already listed (both lists ascending)
    const int ocnt = W.cnt[(size_t)lp * N + p];
    int o = 0;
    int oq = ocnt > 0 ? W.cand[((size_t)lp * S) * N + p] : INT_MAX;
    for (int e = 0; e < k; ++e) {
      const int q = nl[e];
      while (oq < q) {
        ++o;
        oq = o < ocnt ? W.cand[((size_t)lp * S + o) * N + p] : INT_MAX;
      }
      double2 h = make_double2(0.0, 0.0);
      double hr = 0.0;
      if (oq == q) {
        h = W.hxi[((size_t)lp * S + o) * N + p];
        hr = W.hroll[((size_t)lp * S + o) * N + p];
      }
      const size_t at = ((size_t)lq * S + e) * N + p;
      W.cand[at] = q;
      W.hxi[at] = h;
      W.hroll[at] = hr;
    }
    W.cnt[(size_t)lq * N + p] = k;
    W.xb[p] = xp;
  }

  double F0 = 0.0, F1 = 0.0, T = 0.0;
  unsigned touch = 0u;  // candidate slots in contact
  int nbc = 0;          // boundary contacts of this disk
  const bool rolling = P.mu_rolling > 0.0;
  if (live) {
    const int cnt = W.cnt[(size_t)lq * N + p];
    int e0 = 0;
    while (e0 < cnt && W.cand[((size_t)lq * S + e0) * N + p] < p) ++e0;
    // pairs where p is i (partners above p), then where p is j (below)
    for (int pass = 0; pass < 2; ++pass) {
      const int ea = pass ? 0 : e0, eb = pass ? e0 : cnt;
      const bool owner = pass == 0;
      for (int e = ea; e < eb; ++e) {
        const size_t at = ((size_t)lq * S + e) * N + p;
        const int q = W.cand[at];
        const double2 xq = A.x[q];
        const double2 xi_ = owner ? xp : xq, xj_ = owner ? xq : xp;
        const double l0 = min_image(xj_.x - xi_.x, P.period[0]);
        const double l1 = min_image(xj_.y - xi_.y, P.period[1]);
        const double dist = sqrt(l0 * l0 + l1 * l1);
        const double dm = fmax(dist, 1e-30);
        const double nij0 = l0 / dm, nij1 = l1 / dm;
        const double rq = A.radius[q];
        const double ri = owner ? rp : rq, rj = owner ? rq : rp;
        const double ov = (ri + rj) - dist;
        if (!(ov > 0.0)) {
          if (kAdvance) {
            W.hxi[at] = make_double2(0.0, 0.0);
            W.hroll[at] = 0.0;
          }
          if (A.rec && owner) clear_record(A.rec, (size_t)e * N + p);
          continue;
        }
        touch |= 1u << e;
        const double mq = A.mass[q];
        const double mi = owner ? mp : mq, mj = owner ? mq : mp;
        const double2 vq = A.v[q];
        const double omq = A.om[q];
        const double2 vi = owner ? vp : vq, vj = owner ? vq : vp;
        const double omi = owner ? omp : omq, omj = owner ? omq : omp;
        const double n0 = -nij0, n1 = -nij1;
        const double meff = (mi * mj) / (mi + mj);
        const double reff = ri * rj / (ri + rj);
        const double li0 = -ri * n0, li1 = -ri * n1;
        const double lj0 = rj * n0, lj1 = rj * n1;
        const double wi0 = omi * li0, wi1 = omi * li1;
        const double wj0 = omj * lj0, wj1 = omj * lj1;
        const double vi0 = vi.x + -wi1, vi1 = vi.y + wi0;
        const double vj0 = vj.x + -wj1, vj1 = vj.y + wj0;
        double2 h = W.hxi[at];
        double hr = W.hroll[at];
        const LawOut o = contact_law(P, n0, n1, ov, meff, reff, vi0 - vj0, vi1 - vj1,
                                     omi - omj, h.x, h.y, hr);
        if (owner) {
          F0 += o.t0;
          F1 += o.t1;
          T += li0 * o.t1 - li1 * o.t0;
          if (rolling) W.mr[(size_t)e * N + p] = -o.m_r;
          const double vn = (vi0 - vj0) * n0 + (vi1 - vj1) * n1;
          ch[D_DASH] += (o.fn - P.kn * ov) * vn;
          ch[D_SLIDE] += o.slide;
          ch[D_ROLL] += o.rloss;
          ch[D_OV2] += ov * ov;
          ch[D_XI2] += h.x * h.x + h.y * h.y;
          ch[D_KROLL2] += o.kr * (hr * hr);
          ch[D_NPAIR] += 1.0;
          if (A.rec)
            put_record(A.rec, (size_t)e * N + p, (double)p, (double)q, 0.0, n0, n1, l0, l1, o,
                       ov, xi_.x + ri * nij0, xi_.y + ri * nij1);
        } else {
          F0 += -o.t0;
          F1 += -o.t1;
          T += lj0 * -o.t1 - lj1 * -o.t0;
          if (rolling) W.mr[(size_t)e * N + p] = o.m_r;
        }
        if (kAdvance) {
          W.hxi[at] = h;
          W.hroll[at] = hr;
        }
      }
    }
    if (A.rec)
      for (int e = 0; e < S; ++e)
        if (e < e0 || e >= cnt) clear_record(A.rec, (size_t)e * N + p);

    // walls, then hull (engine.py:420-453, 506-521)
    const double4 *wold = W.wst + (size_t)(adv & 1) * ((size_t)N * (P.n_walls + P.n_hull) + 1);
    double4 *wnew = W.wst + (size_t)((adv + 1) & 1) * ((size_t)N * (P.n_walls + P.n_hull) + 1);
    int kw = 0, kh = 0;
    const int nseg = P.n_walls + P.n_hull;
    for (int s = 0; s < nseg; ++s) {
      const bool wall = s < P.n_walls;
      const double *g = wall ? A.walls + 4 * s : A.hull + 4 * (s - P.n_walls);
      SegHit hh;
      const size_t rat = (size_t)(S + s) * N + p;
      if (!seg_hit(g, xp.x, xp.y, rp, hh)) {
        if (A.rec) clear_record(A.rec, rat);
        continue;
      }
      const int rank = wall ? rank_w0 + kw++ : rank_h0 + kh++;
      double4 st = make_double4(0.0, 0.0, 0.0, 0.0);
      if (rank < nst) st = wold[rank];
      const double n0 = hh.n0, n1 = hh.n1;
      const double li0 = -rp * n0, li1 = -rp * n1;
      const double wi0 = omp * li0, wi1 = omp * li1;
      const double vi0 = vp.x + -wi1, vi1 = vp.y + wi0;
      double vj0 = 0.0, vj1 = 0.0;
      if (wall && A.wvel) {
        vj0 = A.wvel[2 * s];
        vj1 = A.wvel[2 * s + 1];
      }
      const LawOut o = contact_law(P, n0, n1, hh.ov, mp, rp, vi0 - vj0, vi1 - vj1, omp - 0.0,
                                   st.x, st.y, st.z);
      F0 += o.t0;
      F1 += o.t1;
      T += li0 * o.t1 - li1 * o.t0;
      if (rolling) W.mr[(size_t)(S + nbc) * N + p] = -o.m_r;
      ++nbc;
      const double vn = (vi0 - vj0) * n0 + (vi1 - vj1) * n1;
      ch[D_DASH] += (o.fn - P.kn * hh.ov) * vn;
      ch[D_SLIDE] += o.slide;
      ch[D_ROLL] += o.rloss;
      ch[D_OV2] += hh.ov * hh.ov;
      ch[D_XI2] += st.x * st.x + st.y * st.y;
      ch[D_KROLL2] += o.kr * (st.z * st.z);
      if (wall) {
        ch[D_WFX] += o.t0;
        ch[D_WFY] += o.t1;
        ch[D_NWALL] += 1.0;
      } else {
        ch[D_HFX] += o.t0;
        ch[D_HFY] += o.t1;
        ch[D_NHULL] += 1.0;
      }
      if (kAdvance) wnew[rank] = st;
      if (A.rec)
        put_record(A.rec, rat, (double)p, -1.0, wall ? 1.0 : 2.0, n0, n1, hh.cp0 - xp.x,
                   hh.cp1 - xp.y, o, hh.ov, hh.cp0, hh.cp1);
    }

    // rolling torques after every lever torque, in the same contact order
    if (rolling) {
      for (int e = e0; e < cnt; ++e)
        if (touch >> e & 1u) T += W.mr[(size_t)e * N + p];
      for (int e = 0; e < e0; ++e)
        if (touch >> e & 1u) T += W.mr[(size_t)e * N + p];
      for (int k = 0; k < nbc; ++k) T += W.mr[(size_t)(S + k) * N + p];
    }
    if (P.use_gravity) {
      F0 = F0 + mp * P.gravity[0];
      F1 = F1 + mp * P.gravity[1];
    }
    if (P.drag_time > 0.0) {
      F0 = F0 + mp * (P.drift[0] - vp.x) / P.drag_time;
      F1 = F1 + mp * (P.drift[1] - vp.y) / P.drag_time;
    }
    if (kAdvance) {
      double v0 = vp.x + P.dt * (F0 / mp), v1 = vp.y + P.dt * (F1 / mp);
      double om = omp + P.dt * (T / A.inertia[p]);
      if (P.use_decay) {
        v0 *= P.decay;
        v1 *= P.decay;
        om *= P.decay;
      }
      double x0 = xp.x + P.dt * v0, x1 = xp.y + P.dt * v1;
      A.theta[p] = A.theta[p] + P.dt * om;
      double xs2[2] = {x0, x1};
#pragma unroll
      for (int a = 0; a < 2; ++a) {
        const double L = P.period[a];
        if (L > 0.0) {
          // numpy remainder: fmod, then shift a remainder of the wrong sign
          double r = fmod(xs2[a], L);
          if (r != 0.0) {
            if ((r < 0.0) != (L < 0.0)) r += L;
          } else {
            r = copysign(0.0, L);
          }
          xs2[a] = r;
        }
      }
      A.xo[p] = make_double2(xs2[0], xs2[1]);
      A.vo[p] = make_double2(v0, v1);
      A.omo[p] = om;
      const double2 xb = W.xb[p];
      const double d0 = min_image(xs2[0] - xb.x, P.period[0]);
      const double d1 = min_image(xs2[1] - xb.y, P.period[1]);
      const double half = 0.5 * P.skin;
      if (!(d0 * d0 + d1 * d1 <= half * half)) atomicOr(&dyn->flag[par ^ 1], 1);
    }
  }

  // ---- ledger partials; the last block reduces them in block order -------
  block_sum<kDch>(ch, red);
  if (threadIdx.x == 0) {
    for (int k = 0; k < kDch; ++k) W.part[(size_t)blockIdx.x * kDch + k] = ch[k];
    __threadfence();
    last = atomicAdd(&dyn->done3, 1u) == gridDim.x - 1;
  }
  __syncthreads();
  if (!last) return;
  __threadfence();
  double acc[kDch];
#pragma unroll
  for (int k = 0; k < kDch; ++k) acc[k] = 0.0;
  for (int b = threadIdx.x; b < (int)gridDim.x; b += kThreads)
#pragma unroll
    for (int k = 0; k < kDch; ++k) acc[k] += atomic_load_f64(&W.part[(size_t)b * kDch + k]);
  __syncthreads();
  block_sum<kDch>(acc, red);
  if (threadIdx.x == 0) {
    double *out = A.out;
    if (kAdvance && A.ledger) {
      out[TDM_DISK_NORMAL_DASHPOT] = out[TDM_DISK_NORMAL_DASHPOT] + acc[D_DASH] * P.dt;
      out[TDM_DISK_FRICTION] = out[TDM_DISK_FRICTION] - acc[D_SLIDE];
      out[TDM_DISK_ROLLING] = out[TDM_DISK_ROLLING] - acc[D_ROLL];
    }
    out[TDM_DISK_ELASTIC] = 0.5 * P.kn * acc[D_OV2];
    out[TDM_DISK_ELASTIC + 1] = 0.5 * P.kt * acc[D_XI2];
    out[TDM_DISK_ELASTIC + 2] = rolling ? 0.5 * acc[D_KROLL2] : 0.0;
    out[TDM_DISK_WALL_FORCE] = -acc[D_WFX];
    out[TDM_DISK_WALL_FORCE + 1] = -acc[D_WFY];
    out[TDM_DISK_HULL_FORCE] = -acc[D_HFX];
    out[TDM_DISK_HULL_FORCE + 1] = -acc[D_HFY];
    out[TDM_DISK_CONTACTS] = acc[D_NPAIR];
    out[TDM_DISK_CONTACTS + 1] = acc[D_NWALL];
    out[TDM_DISK_CONTACTS + 2] = acc[D_NHULL];
    if (rebuild) {
      dyn->lp = lp ^ 1;
      dyn->cellp = cellp ^ 1;
      dyn->rebuilds += 1;
    }
    dyn->flag[par] = 0;
    dyn->launch = launch + 1;
    if (kAdvance) {
      dyn->adv = adv + 1;
      dyn->nst = dyn->nw + dyn->nh;
    }
    // other blocks OR into status during this launch: read it from L2
    out[TDM_DISK_STATUS] = (double)atomic_load_i32(&dyn->status);
    out[TDM_DISK_REBUILDS] = (double)dyn->rebuilds;
    dyn->done3 = 0u;
  }
}

__global__ void disk_init_kernel(DiskDyn *dyn)
{
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    DiskDyn d = {};
    d.flag[0] = 1;
    *dyn = d;
  }
}

__global__ void disk_mark_kernel(DiskDyn *dyn)
{
  if (threadIdx.x == 0 && blockIdx.x == 0) dyn->flag[dyn->launch & 1] = 1;
}

int check_disk(const tdm_disk_arrays *a, const tdm_disk_params *p, DiskWs *w)
{
  if (!a || !p) return fail(TDM_EINVAL, "null disk arrays or params");
  if (p->n <= 0 || p->nblocks != tdm_disk_blocks(p->n))
    return fail(TDM_EINVAL, "disk count / nblocks mismatch");
  if (p->cand_slots <= 0 || p->cand_slots > kDiskCap)
    return fail(TDM_EINVAL, "cand_slots must be in 1..24");
  if (p->cell_cap <= 0 || p->ncell[0] <= 0 || p->ncell[1] <= 0)
    return fail(TDM_EINVAL, "bad cell grid");
  if (p->n_walls < 0 || p->n_walls > 255 || p->n_hull < 0 || p->n_hull > 255)
    return fail(TDM_EINVAL, "at most 255 wall and 255 hull segments");
  if ((size_t)p->ncell[0] * p->ncell[1] > (size_t)INT_MAX / (size_t)p->cell_cap)
    return fail(TDM_EINVAL, "cell grid too large");
  // a periodic axis with one cell holds every disk in it, whatever its edge
  for (int ax = 0; ax < 2; ++ax)
    if (!(p->period[ax] > 0.0 && p->ncell[ax] == 1) && !(p->cell_h[ax] >= p->reach))
      return fail(TDM_EINVAL, "cell edge must be at least the candidate reach");
  if (!(p->reach > 0.0)) return fail(TDM_EINVAL, "candidate reach must be positive");
  if (!a->x || !a->v || !a->omega || !a->x_alt || !a->v_alt || !a->omega_alt || !a->theta ||
      !a->radius || !a->mass || !a->inertia || !a->out || !a->workspace)
    return fail(TDM_EINVAL, "null disk array");
  if ((p->n_walls && !a->walls) || (p->n_hull && !a->hull))
    return fail(TDM_EINVAL, "null segment array");
  if (reinterpret_cast<uintptr_t>(a->workspace) & 255)
    return fail(TDM_EINVAL, "workspace must be 256-byte aligned");
  size_t total = 0;
  carve(static_cast<char *>(a->workspace), p->n, p->ncell[0] * p->ncell[1], p->cell_cap,
        p->cand_slots, p->n_walls, p->n_hull, w, &total);
  if (a->workspace_bytes < total) return fail(TDM_ESPACE, "disk workspace too small");
  return TDM_OK;
}

DiskArgs disk_args(const tdm_disk_arrays *a, const DiskWs &w, int parity, int ledger,
                   bool records)
{
  DiskArgs A;
  const bool odd = parity & 1;
  A.x = reinterpret_cast<const double2 *>(odd ? a->x_alt : a->x);
  A.v = reinterpret_cast<const double2 *>(odd ? a->v_alt : a->v);
  A.om = odd ? a->omega_alt : a->omega;
  A.xo = reinterpret_cast<double2 *>(odd ? a->x : a->x_alt);
  A.vo = reinterpret_cast<double2 *>(odd ? a->v : a->v_alt);
  A.omo = odd ? a->omega : a->omega_alt;
  A.theta = a->theta;
  A.radius = a->radius;
  A.mass = a->mass;
  A.inertia = a->inertia;
  A.walls = a->walls;
  A.hull = a->hull;
  A.wvel = a->wall_velocity;
  A.out = a->out;
  A.rec = records ? a->records : nullptr;
  A.w = w;
  A.ledger = ledger;
  return A;
}

int launch_disk_steps(const tdm_disk_arrays *a, const tdm_disk_params *p, const DiskWs &w,
                      int nsteps, int flags, hipStream_t s, hipEvent_t *ev = nullptr)
{
  const int nb = p->nblocks;
  for (int i = 0; i < nsteps; ++i) {
    const bool rec = (flags & TDM_DISK_RECORD) && a->records && i == nsteps - 1;
    const DiskArgs A = disk_args(a, w, i, flags & TDM_DISK_LEDGER, rec);
    if (ev) HIP_TRY(hipEventRecord(ev[4 * i], s));
    disk_prep_kernel<<<nb, kThreads, 0, s>>>(*p, A);
    if (ev) HIP_TRY(hipEventRecord(ev[4 * i + 1], s));
    if (ev) HIP_TRY(hipEventRecord(ev[4 * i + 2], s));
    disk_step_kernel<true><<<nb, kThreads, 0, s>>>(*p, A);
    if (ev) HIP_TRY(hipEventRecord(ev[4 * i + 3], s));
  }
  HIP_TRY(hipGetLastError());
  return TDM_OK;
}

}  // namespace
}  // namespace tdm

using namespace tdm;

extern "C" {

int tdm_disk_abi_sizes(size_t *params_bytes, size_t *arrays_bytes)
{
  if (!params_bytes || !arrays_bytes) return fail(TDM_EINVAL, "null output");
  *params_bytes = sizeof(tdm_disk_params);
  *arrays_bytes = sizeof(tdm_disk_arrays);
  return TDM_OK;
}

int tdm_disk_blocks(int n) { return n <= 0 ? 0 : (n + kThreads - 1) / kThreads; }

size_t tdm_disk_workspace_bytes(int n, int ncells, int cell_cap, int cand_slots, int n_walls,
                                int n_hull)
{
  DiskWs w;
  size_t total = 0;
  carve(nullptr, n, ncells, cell_cap, cand_slots, n_walls, n_hull, &w, &total);
  return total;
}

int tdm_disk_setup(const tdm_disk_arrays *a, const tdm_disk_params *p, void *stream)
{
  DiskWs w;
  int rc = check_disk(a, p, &w);
  if (rc) return rc;
  hipStream_t s = static_cast<hipStream_t>(stream);
  const size_t ncells = (size_t)p->ncell[0] * p->ncell[1];
  fill_kernel<int><<<256, 256, 0, s>>>(w.cnt, 2 * (size_t)p->n, 0);
  fill_kernel<int><<<256, 256, 0, s>>>(w.ccount, 2 * ncells, 0);
  disk_init_kernel<<<1, 64, 0, s>>>(w.dyn);
  HIP_TRY(hipGetLastError());
  return TDM_OK;
}

int tdm_disk_rebin(const tdm_disk_arrays *a, const tdm_disk_params *p, void *stream)
{
  DiskWs w;
  int rc = check_disk(a, p, &w);
  if (rc) return rc;
  disk_mark_kernel<<<1, 64, 0, static_cast<hipStream_t>(stream)>>>(w.dyn);
  HIP_TRY(hipGetLastError());
  return TDM_OK;
}

int tdm_disk_step(const tdm_disk_arrays *a, const tdm_disk_params *p, int nsteps, int flags,
                  void *stream)
{
  DiskWs w;
  int rc = check_disk(a, p, &w);
  if (rc) return rc;
  if (nsteps < 0) return fail(TDM_EINVAL, "nsteps must be nonnegative");
  return launch_disk_steps(a, p, w, nsteps, flags, static_cast<hipStream_t>(stream));
}

int tdm_disk_inspect(const tdm_disk_arrays *a, const tdm_disk_params *p, void *stream)
{
  DiskWs w;
  int rc = check_disk(a, p, &w);
  if (rc) return rc;
  hipStream_t s = static_cast<hipStream_t>(stream);
  const DiskArgs A = disk_args(a, w, 0, 0, a->records != nullptr);
  disk_prep_kernel<<<p->nblocks, kThreads, 0, s>>>(*p, A);
  disk_step_kernel<false><<<p->nblocks, kThreads, 0, s>>>(*p, A);
  HIP_TRY(hipGetLastError());
  return TDM_OK;
}

int tdm_disk_step_timed(const tdm_disk_arrays *a, const tdm_disk_params *p, int nsteps,
                        void *stream, double *mean_step_us, double *mean_prep_us)
{
  DiskWs w;
  int rc = check_disk(a, p, &w);
  if (rc) return rc;
  if (nsteps <= 0 || (nsteps & 1) || !mean_step_us || !mean_prep_us)
    return fail(TDM_EINVAL, "bad timing request");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int nev = 4 * nsteps;
  hipEvent_t *ev = new hipEvent_t[nev]();
  int made = 0;
  for (; made < nev; ++made)
    if (hipEventCreate(&ev[made]) != hipSuccess) break;
  if (made < nev) rc = fail(TDM_EHIP, "hipEventCreate failed");
  if (rc == TDM_OK) rc = launch_disk_steps(a, p, w, nsteps, 0, s, ev);
  double tstep = 0.0, tprep = 0.0;
  if (rc == TDM_OK) {
    hipError_t e = hipStreamSynchronize(s);
    for (int i = 0; i < nsteps && e == hipSuccess; ++i) {
      float ms = 0.0f;
      e = hipEventElapsedTime(&ms, ev[4 * i], ev[4 * i + 1]);
      tprep += ms;
      if (e == hipSuccess) e = hipEventElapsedTime(&ms, ev[4 * i + 2], ev[4 * i + 3]);
      tstep += ms;
    }
    if (e != hipSuccess) rc = fail(TDM_EHIP, hipGetErrorString(e));
  }
  for (int i = 0; i < made; ++i) (void)hipEventDestroy(ev[i]);
  delete[] ev;
  if (rc == TDM_OK) {
    *mean_step_us = tstep * 1e3 / nsteps;
    *mean_prep_us = tprep * 1e3 / nsteps;
  }
  return rc;
}

int tdm_disk_graph_create(const tdm_disk_arrays *a, const tdm_disk_params *p, int pairs,
                          int flags, void *stream, tdm_graph **out)
{
  DiskWs w;
  int rc = check_disk(a, p, &w);
  if (rc) return rc;
  if (!out || pairs <= 0) return fail(TDM_EINVAL, "bad graph request");
  hipStream_t s = static_cast<hipStream_t>(stream);
  HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  rc = launch_disk_steps(a, p, w, 2 * pairs, flags & TDM_DISK_LEDGER, s);
  hipGraph_t g;
  hipError_t e = hipStreamEndCapture(s, &g);
  if (rc) return rc;
  if (e != hipSuccess) return fail(TDM_EHIP, hipGetErrorString(e));
  hipGraphExec_t ex;
  e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    (void)hipGraphDestroy(g);
    return fail(TDM_EHIP, hipGetErrorString(e));
  }
  *out = new tdm_graph{g, ex};
  return TDM_OK;
}

}  // extern "C"
