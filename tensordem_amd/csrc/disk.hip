// Rigid frictional disks with rotation: the step of the reference's
// DiskSystem (src/tensordem/granular2d/engine.py:399-596) on the device.
//
// Two kernels per step, one thread per disk:
//
//   disk_prep_kernel   counts each disk's wall and hull contacts at the
//                      current positions, per block; when a candidate search
//                      is due it also drops every disk into the cell table.
//                      Its first block reduces the previous step kernel's
//                      per-block ledger partials in block order.
//   disk_step_kernel   (search step: builds the disk's sorted candidate list
//                      from the 3x3 cells, carrying each surviving partner's
//                      spring over), evaluates every contact of the disk in
//                      the reference's order, integrates, and writes the new
//                      spring state.  A boundary contact's spring is stored at
//                      its rank among all boundary contacts (how the reference
//                      keys them, engine.py:452-456): each block adds up the
//                      counts of the blocks before it.
//
// No kernel waits on another block: per-launch bookkeeping is indexed by the
// launch parity, which the host passes in, so nothing needs a device-wide
// fence (on a chip of several L2 caches an agent-scope release per block
// costs more than the step).  A flush kernel reduces the last step's partials
// at the end of each call.
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

constexpr int kDiskCap = 24;  // candidate slots per disk (bits of the contact mask)
constexpr int kDch = 14;      // ledger partial channels
constexpr int kBatch = 4;     // candidates fetched together
constexpr int kPrepBlocks = 512;
constexpr int kParamSegs = 8;     // segment boxes carried in tdm_disk_params.seg_box  // grid of the prep kernel (grid-stride binning)
enum {
  D_DASH, D_SLIDE, D_ROLL, D_OV2, D_XI2, D_KROLL2, D_WFX, D_WFY, D_HFX, D_HFY, D_NPAIR,
  D_NWALL, D_NHULL, D_NCAND
};

// device-side step state (in the workspace); [k] = launch or step parity k
struct DiskDyn {
  int lpar[2];   // list buffer holding the candidate lists at that launch parity
  int cpar[2];   // cell table a search at that launch parity bins into
  int flag[2];   // candidate search due at the launch of that parity
  int nst[2];    // boundary contacts stored in wst[k] (step parity)
  int rebuilds;
  int status;
  int pend;      // partials of the last step kernel: 0 none, else 1 | 2 (stepped) | 4 (ledger)
  int tot[2];    // wall and hull contacts at the current positions
  int pad[3];
};

struct DiskWs {
  DiskDyn *dyn;
  int *cand;          // [2][S][N]
  double2 *hxi;       // [2][S][N] spring of each candidate pair
  double *hroll;      // [2][S][N] rolling angle of each candidate pair
  int *cnt;           // [2][N] candidates per disk | (those below the disk's index) << 8
  unsigned *cmask;    // [2][N] bit e: slot e was in contact at the last step (spring stored)
  double2 *xb;        // [N] position at the last search
  int *ccount;        // [2][ncells]
  int *centry;        // [2][ncells][cell_cap]
  double4 *wst;       // [2][wcap] boundary springs by rank: xi, roll
  int *blkcnt;        // [2][nblocks] packed wall | hull << 16 contacts per block, by step parity
  int2 *blkoff;       // [nblocks] exclusive wall, hull offsets of the blocks
  double *part;       // [nblocks][kDch]
  double *mr;         // [S + n_walls + n_hull][N] rolling torque terms
  double *segbox;     // [n_walls + n_hull][4] bounding boxes of the segments
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
  w->cmask = reinterpret_cast<unsigned *>(take(sizeof(unsigned) * 2 * (size_t)n));
  w->xb = reinterpret_cast<double2 *>(take(sizeof(double2) * (size_t)n));
  w->ccount = reinterpret_cast<int *>(take(sizeof(int) * 2 * (size_t)ncells));
  w->centry = reinterpret_cast<int *>(take(sizeof(int) * 2 * (size_t)ncells * cap));
  w->wst = reinterpret_cast<double4 *>(take(sizeof(double4) * 2 * wcap));
  w->blkcnt = reinterpret_cast<int *>(take(sizeof(int) * 2 * (size_t)nb));
  w->blkoff = reinterpret_cast<int2 *>(take(sizeof(int2) * (size_t)nb));
  w->part = reinterpret_cast<double *>(take(sizeof(double) * kDch * (size_t)nb));  // [kDch][nb]
  w->mr = reinterpret_cast<double *>(take(sizeof(double) * (size_t)(S + nw + nh) * n));
  w->segbox = reinterpret_cast<double *>(take(sizeof(double) * 4 * (size_t)(nw + nh + 1)));
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
  int par;    // launch parity
  int spar;   // step parity (boundary spring buffers)
  int mode;   // 1 evaluation, | 2 stepped, | 4 ledger
};

__device__ inline double min_image(double d, double L) { return L > 0.0 ? d - L * rint(d / L) : d; }

// np.maximum(0.0, x) and np.maximum(x, 0.0): NaN propagates, ties keep 0.0
__device__ inline double max0_first(double x) { return (0.0 >= x) ? 0.0 : x; }
__device__ inline double max0_second(double x) { return (x >= 0.0 || x != x) ? x : 0.0; }

// Closest point of a segment (engine.py:352-375): contact when the overlap
// r - |d| is positive and |d| > 1e-12; n points from the segment to the disk.
struct SegHit {
  double n0, n1, ov, cp0, cp1;
};

// Quick reject: a disk farther than r (with a 1e-9 relative margin, far above
// the rounding of the exact test) from the segment's bounding box `box`
// (min x, min y, max x, max y) cannot touch it.
__device__ inline bool seg_near(const double *box, double px, double py, double r)
{
  const double m = r * (1.0 + 1e-9);
  return !(box[0] - px > m || px - box[2] > m || box[1] - py > m || py - box[3] > m);
}

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

__device__ inline const double *segment(const tdm_disk_params &P, const DiskArgs &A, int s)
{
  return s < P.n_walls ? A.walls + 4 * s : A.hull + 4 * (s - P.n_walls);
}

// Segments the disk touches: bit s of the mask (first 64 segments) and the
// wall | hull << 16 counts.
// Segments the disk touches: bit s of the mask (first 64 segments) and the
// wall | hull << 16 counts.  The first kParamSegs boxes come with the launch
// parameters (the loop over them is unrolled, so they stay in scalar
// registers), the rest from the workspace.
__device__ inline int boundary_hits(const tdm_disk_params &P, const DiskArgs &A, double px,
                                    double py, double r, unsigned long long &mask)
{
  int cw = 0, chh = 0;
  mask = 0ull;
  SegHit h;
  const int nseg = P.n_walls + P.n_hull;
#pragma unroll
  for (int s = 0; s < kParamSegs; ++s) {
    if (s >= nseg) break;
    if (!seg_near(P.seg_box[s], px, py, r)) continue;
    if (!seg_hit(segment(P, A, s), px, py, r, h)) continue;
    mask |= 1ull << s;
    if (s < P.n_walls) ++cw;
    else ++chh;
  }
  for (int s = kParamSegs; s < nseg; ++s) {
    if (!seg_near(A.w.segbox + 4 * s, px, py, r)) continue;
    if (!seg_hit(segment(P, A, s), px, py, r, h)) continue;
    if (s < 64) mask |= 1ull << s;
    if (s < P.n_walls) ++cw;
    else ++chh;
  }
  return cw | (chh << 16);
}

__device__ inline int boundary_counts(const tdm_disk_params &P, const DiskArgs &A, double px,
                                      double py, double r)
{
  unsigned long long mask;
  return boundary_hits(P, A, px, py, r, mask);
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

// Disks [256 lb, 256 lb + 256) of logical block lb.  Hardware blocks go to
// the 8 XCDs round robin; this mapping gives each XCD a contiguous run of
// logical blocks, so the partners of a disk (neighbouring indices and, in a
// spatially ordered assembly, the next rows) are mostly read through the
// same L2.
__device__ inline int logical_block()
{
  const int nb = gridDim.x;
  if (nb & 7) return blockIdx.x;
  return (blockIdx.x & 7) * (nb >> 3) + (blockIdx.x >> 3);
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

// One contact of the law (engine.py:458-537).  (n0, n1) points from j to
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

// Reduces the pending per-block partials of the last step kernel into
// `out` in block order (one block; deterministic).
__device__ void reduce_pending(const tdm_disk_params &P, const DiskArgs &A, double (*red)[kCh])
{
  DiskDyn *dyn = A.w.dyn;
  const int pend = dyn->pend;
  if (!pend) return;
  double acc[kDch];
#pragma unroll
  for (int k = 0; k < kDch; ++k) acc[k] = 0.0;
  // channel-major partials: consecutive threads read consecutive blocks
#pragma unroll 4
  for (int b = threadIdx.x; b < P.nblocks; b += kThreads)
#pragma unroll
    for (int k = 0; k < kDch; ++k) acc[k] += A.w.part[(size_t)k * P.nblocks + b];
  block_sum<kDch>(acc, red);
  if (threadIdx.x == 0) {
    double *out = A.out;
    if (pend & 4) {
      out[TDM_DISK_NORMAL_DASHPOT] = out[TDM_DISK_NORMAL_DASHPOT] + acc[D_DASH] * P.dt;
      out[TDM_DISK_FRICTION] = out[TDM_DISK_FRICTION] - acc[D_SLIDE];
      out[TDM_DISK_ROLLING] = out[TDM_DISK_ROLLING] - acc[D_ROLL];
    }
    out[TDM_DISK_ELASTIC] = 0.5 * P.kn * acc[D_OV2];
    out[TDM_DISK_ELASTIC + 1] = 0.5 * P.kt * acc[D_XI2];
    out[TDM_DISK_ELASTIC + 2] = P.mu_rolling > 0.0 ? 0.5 * acc[D_KROLL2] : 0.0;
    out[TDM_DISK_WALL_FORCE] = -acc[D_WFX];
    out[TDM_DISK_WALL_FORCE + 1] = -acc[D_WFY];
    out[TDM_DISK_HULL_FORCE] = -acc[D_HFX];
    out[TDM_DISK_HULL_FORCE + 1] = -acc[D_HFY];
    out[TDM_DISK_CONTACTS] = acc[D_NPAIR];
    out[TDM_DISK_CONTACTS + 1] = acc[D_NWALL];
    out[TDM_DISK_CONTACTS + 2] = acc[D_NHULL];
    out[TDM_DISK_STATUS] = (double)dyn->status;
    out[TDM_DISK_REBUILDS] = (double)dyn->rebuilds;
    out[TDM_DISK_CANDIDATES] = acc[D_NCAND];
    dyn->pend = 0;
  }
}

// Exclusive offsets of the per-block wall and hull contact counts (one block).
__device__ void scan_blocks(const tdm_disk_params &P, const DiskArgs &A)
{
  __shared__ int sw_s[kThreads / 64], sh_s[kThreads / 64];
  const int nb = P.nblocks;
  const int per = (nb + kThreads - 1) / kThreads;
  const int b0 = min(nb, (int)threadIdx.x * per), b1 = min(nb, b0 + per);
  const int *cnt = A.w.blkcnt + (size_t)A.spar * nb;
  int sw = 0, sh = 0;
#pragma unroll 16
  for (int b = b0; b < b1; ++b) {
    const int v = cnt[b];
    sw += v & 0xFFFF;
    sh += v >> 16;
  }
  // inclusive scan: inside each wave by shuffles, then the wave totals
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int iw = sw, ih = sh;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const int tw = __shfl_up(iw, d), th = __shfl_up(ih, d);
    if (lane >= d) {
      iw += tw;
      ih += th;
    }
  }
  if (lane == 63) {
    sw_s[wave] = iw;
    sh_s[wave] = ih;
  }
  __syncthreads();
  int bw = 0, bh = 0, tw = 0, th = 0;
#pragma unroll
  for (int k = 0; k < kThreads / 64; ++k) {
    if (k < wave) {
      bw += sw_s[k];
      bh += sh_s[k];
    }
    tw += sw_s[k];
    th += sh_s[k];
  }
  int ow = bw + iw - sw, oh = bh + ih - sh;
#pragma unroll 16
  for (int b = b0; b < b1; ++b) {
    const int v = cnt[b];
    A.w.blkoff[b] = make_int2(ow, oh);
    ow += v & 0xFFFF;
    oh += v >> 16;
  }
  if (threadIdx.x == 0) {
    A.w.dyn->tot[0] = tw;
    A.w.dyn->tot[1] = th;
  }
}

// Per-block wall and hull contact counts of the disks at A.x (start of a
// call; later steps get them from the step kernel, after integrating).
__global__ __launch_bounds__(kThreads) void disk_count_kernel(tdm_disk_params P, DiskArgs A)
{
  __shared__ int bsum;
  const int lb = logical_block();
  const int p = lb * kThreads + threadIdx.x;
  if (threadIdx.x == 0) bsum = 0;
  __syncthreads();
  if (p < P.n) {
    const double2 xp = A.x[p];
    const int packed = boundary_counts(P, A, xp.x, xp.y, A.radius[p]);
    if (packed) atomicAdd(&bsum, packed);
  }
  __syncthreads();
  if (threadIdx.x == 0) A.w.blkcnt[(size_t)A.spar * P.nblocks + lb] = bsum;
}

__global__ __launch_bounds__(kThreads) void disk_prep_kernel(tdm_disk_params P, DiskArgs A)
{
  __shared__ double red[4][kCh];
  DiskDyn *dyn = A.w.dyn;
  const int p = blockIdx.x * kThreads + threadIdx.x;
  const bool rebuild = dyn->flag[A.par] != 0;
  const int cellp = dyn->cpar[A.par];
  if (P.n_walls + P.n_hull) {
    // the step kernel of this launch adds the next positions' counts here
    int *next = A.w.blkcnt + (size_t)(A.spar ^ 1) * P.nblocks;
    for (int b = p; b < P.nblocks; b += gridDim.x * kThreads) next[b] = 0;
  }
  if (blockIdx.x == 0) {
    reduce_pending(P, A, red);
    if (P.n_walls + P.n_hull) {
      __syncthreads();
      scan_blocks(P, A);
    }
    // the step kernel of this launch raises the flag for the next one
    if (threadIdx.x == 0) dyn->flag[A.par ^ 1] = 0;
  }
  if (!rebuild) return;
  // a small grid: most launches only do block 0's work
  for (int q = p; q < P.n; q += gridDim.x * kThreads) {
    const double2 xq = A.x[q];
    const int ncells = P.ncell[0] * P.ncell[1];
    const int c = cell_coord(P, 0, xq.x) + P.ncell[0] * cell_coord(P, 1, xq.y);
    const size_t cc = (size_t)cellp * ncells + c;
    const int slot = atomicAdd(&A.w.ccount[cc], 1);
    if (slot < P.cell_cap) A.w.centry[cc * P.cell_cap + slot] = q;
    else atomicOr(&dyn->status, TDM_ST_CELL_OVERFLOW);
  }
}

__global__ __launch_bounds__(kThreads) void disk_flush_kernel(tdm_disk_params P, DiskArgs A)
{
  __shared__ double red[4][kCh];
  reduce_pending(P, A, red);
}

// kSums: accumulate the ledger, energy, reaction and count partials (steps
// whose sums nobody reads skip them)
template <bool kAdvance, bool kSums>
__global__ __launch_bounds__(kThreads) void disk_step_kernel(tdm_disk_params P, DiskArgs A)
{
  __shared__ double red[4][kCh];
  __shared__ int scan[kThreads];
  DiskDyn *dyn = A.w.dyn;
  const DiskWs &W = A.w;
  const int N = P.n, S = P.cand_slots;
  const int lb = logical_block();
  const int p = lb * kThreads + threadIdx.x;
  const bool live = p < N;
  const int par = A.par;
  const bool rebuild = dyn->flag[par] != 0;
  const int lp = dyn->lpar[par];
  const int lq = rebuild ? lp ^ 1 : lp;
  const int cellp = dyn->cpar[par];
  const int nst = dyn->nst[A.spar];
  const int ncells = P.ncell[0] * P.ncell[1];
  const bool bounded = P.n_walls + P.n_hull > 0;
  double ch[kDch];
#pragma unroll
  for (int k = 0; k < kDch; ++k) ch[k] = 0.0;

  if (rebuild) {  // clear the idle cell table for the next search
    int *cc = W.ccount + (size_t)(cellp ^ 1) * ncells;
    for (int c = blockIdx.x * kThreads + threadIdx.x; c < ncells; c += gridDim.x * kThreads)
      cc[c] = 0;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    // bookkeeping read only by the next launch (the other parity)
    dyn->lpar[par ^ 1] = lq;
    dyn->cpar[par ^ 1] = rebuild ? cellp ^ 1 : cellp;
    if (rebuild) dyn->rebuilds += 1;
    if (kSums) dyn->pend = A.mode;
  }

  double2 xp = make_double2(0.0, 0.0), vp = xp;
  double omp = 0.0, rp = 0.0, mp = 1.0;
  int packed = 0;
  unsigned long long smask = 0ull;
  if (live) {
    xp = A.x[p];
    vp = A.v[p];
    omp = A.om[p];
    rp = A.radius[p];
    mp = A.mass[p];
    if (bounded) packed = boundary_hits(P, A, xp.x, xp.y, rp, smask);
  }
  // boundary-contact ranks: the block's offsets, then the prefix inside it
  // (blocks without boundary contacts, known from the offsets, skip it)
  int rank_w0 = 0, rank_h0 = 0;
  int2 boff = make_int2(0, 0), bnext = boff;
  if (bounded) {
    boff = W.blkoff[lb];
    bnext = lb + 1 < (int)gridDim.x ? W.blkoff[lb + 1] : make_int2(dyn->tot[0], dyn->tot[1]);
  }
  if (bounded && (bnext.x - boff.x) + (bnext.y - boff.y) > 0) {
    scan[threadIdx.x] = packed;
    __syncthreads();
    for (int d = 1; d < kThreads; d <<= 1) {
      const int a = threadIdx.x >= d ? scan[threadIdx.x - d] : 0;
      __syncthreads();
      scan[threadIdx.x] += a;
      __syncthreads();
    }
    const int excl = scan[threadIdx.x] - packed;
    rank_w0 = boff.x + (excl & 0xFFFF);
    rank_h0 = dyn->tot[0] + boff.y + (excl >> 16);
  }
  if (kAdvance && blockIdx.x == 0 && threadIdx.x == 0)
    dyn->nst[A.spar ^ 1] = bounded ? dyn->tot[0] + dyn->tot[1] : 0;

  // ---- candidate search (engine.py:298-350 gives the pair set) -------------
  if (live && rebuild) {
    // the list is kept sorted in place in the disk's own slots (cached)
    int *cl = W.cand + (size_t)lq * S * N + p;
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
          for (; e > 0; --e) {
            const int prev = cl[(size_t)(e - 1) * N];
            if (prev < q) break;
            cl[(size_t)e * N] = prev;
          }
          cl[(size_t)e * N] = q;
        }
      }
    // carry the springs of partners in contact at the last step (both
    // lists ascending); other slots need no spring
    const int ocnt = W.cnt[(size_t)lp * N + p] & 0xFF;
    const unsigned omask = W.cmask[(size_t)lp * N + p];
    unsigned nmask = 0u;
    int below = 0;
    int o = 0;
    int oq = ocnt > 0 ? W.cand[((size_t)lp * S) * N + p] : INT_MAX;
    for (int e = 0; e < k; ++e) {
      const int q = cl[(size_t)e * N];
      below += q < p;
      while (oq < q) {
        ++o;
        oq = o < ocnt ? W.cand[((size_t)lp * S + o) * N + p] : INT_MAX;
      }
      const size_t at = ((size_t)lq * S + e) * N + p;
      if (oq == q && (omask >> o & 1u)) {
        W.hxi[at] = W.hxi[((size_t)lp * S + o) * N + p];
        W.hroll[at] = W.hroll[((size_t)lp * S + o) * N + p];
        nmask |= 1u << e;
      }
    }
    W.cnt[(size_t)lq * N + p] = k | (below << 8);
    W.cmask[(size_t)lq * N + p] = nmask;
    W.xb[p] = xp;
  }

  double F0 = 0.0, F1 = 0.0, T = 0.0;
  int packed_new = 0;
  unsigned touch = 0u;  // candidate slots in contact
  // rolling torques, added after every lever torque in contact order: the
  // first four in registers, later ones in the disk's scratch column
  int nroll = 0;
  double rr0 = 0.0, rr1 = 0.0, rr2 = 0.0, rr3 = 0.0;
  auto push_roll = [&](double v) {
    if (nroll == 0) rr0 = v;
    else if (nroll == 1) rr1 = v;
    else if (nroll == 2) rr2 = v;
    else if (nroll == 3) rr3 = v;
    else W.mr[(size_t)(nroll - 4) * N + p] = v;
    ++nroll;
  };
  const bool rolling = P.mu_rolling > 0.0;
  if (live) {
    const int cw = W.cnt[(size_t)lq * N + p];
    const int cnt = cw & 0xFF, e0 = cw >> 8;
    const unsigned mk = W.cmask[(size_t)lq * N + p];
    if (kSums) ch[D_NCAND] = (double)cnt;
    const int *cl = W.cand + (size_t)lq * S * N + p;
    // Phase 1, every candidate: does it overlap?  The test uses the same
    // operands as the evaluation (|x_q - x_p| is the canonical |x_j - x_i|
    // and r_p + r_q the canonical r_i + r_j), with a squared-distance
    // pre-test whose 1e-12 relative margin lies far above the rounding of the
    // exact test.  kBatch candidates are fetched together.  Bit j of `hit` is
    // candidate j in evaluation order: partners above p (slots e0..cnt-1,
    // where p is i), then below (slots 0..e0-1, where p is j).
    const int nup = cnt - e0;
    unsigned hit = 0u;
    for (int j0 = 0; j0 < cnt; j0 += kBatch) {
      int qb[kBatch];
      double2 xb_[kBatch];
      double rb[kBatch];
#pragma unroll
      for (int u = 0; u < kBatch; ++u) {
        const int j = j0 + u;
        const int e = j < nup ? e0 + j : j - nup;
        qb[u] = j < cnt ? cl[(size_t)e * N] : p;
      }
#pragma unroll
      for (int u = 0; u < kBatch; ++u) {
        xb_[u] = A.x[qb[u]];
        rb[u] = A.radius[qb[u]];
      }
#pragma unroll
      for (int u = 0; u < kBatch; ++u) {
        const int j = j0 + u;
        const double l0 = min_image(xb_[u].x - xp.x, P.period[0]);
        const double l1 = min_image(xb_[u].y - xp.y, P.period[1]);
        const double l2 = l0 * l0 + l1 * l1;
        const double rs = rp + rb[u];
        if (j < cnt && !(l2 > rs * rs * (1.0 + 1e-12)) && rs - sqrt(l2) > 0.0)
          hit |= 1u << j;
      }
    }
    // Phase 2, contacts only, in evaluation order (a wave runs the law as
    // often as its busiest disk has contacts, not once per candidate)
    while (hit) {
      const int j = __builtin_ctz(hit);
      hit &= hit - 1u;
      {
        const int e = j < nup ? e0 + j : j - nup;
        const bool owner = j < nup;
        touch |= 1u << e;
        const size_t at = ((size_t)lq * S + e) * N + p;
        const int q = cl[(size_t)e * N];
        const double2 xq = A.x[q];
        const double rq = A.radius[q];
        const double2 xi_ = owner ? xp : xq, xj_ = owner ? xq : xp;
        const double l0 = min_image(xj_.x - xi_.x, P.period[0]);
        const double l1 = min_image(xj_.y - xi_.y, P.period[1]);
        const double ri = owner ? rp : rq, rj = owner ? rq : rp;
        const double dist = sqrt(l0 * l0 + l1 * l1);
        const double ov = (ri + rj) - dist;
        const double dm = fmax(dist, 1e-30);
        const double nij0 = l0 / dm, nij1 = l1 / dm;
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
        double2 h = make_double2(0.0, 0.0);
        double hr = 0.0;
        if (mk >> e & 1u) {
          h = W.hxi[at];
          hr = W.hroll[at];
        }
        const LawOut o = contact_law(P, n0, n1, ov, meff, reff, vi0 - vj0, vi1 - vj1,
                                     omi - omj, h.x, h.y, hr);
        if (owner) {
          F0 += o.t0;
          F1 += o.t1;
          T += li0 * o.t1 - li1 * o.t0;
          if (rolling) push_roll(-o.m_r);
          if (kSums) {
            const double vn = (vi0 - vj0) * n0 + (vi1 - vj1) * n1;
            ch[D_DASH] += (o.fn - P.kn * ov) * vn;
            ch[D_SLIDE] += o.slide;
            ch[D_ROLL] += o.rloss;
            ch[D_OV2] += ov * ov;
            ch[D_XI2] += h.x * h.x + h.y * h.y;
            ch[D_KROLL2] += o.kr * (hr * hr);
            ch[D_NPAIR] += 1.0;
          }
          if (A.rec)
            put_record(A.rec, (size_t)e * N + p, (double)p, (double)q, 0.0, n0, n1, l0, l1, o,
                       ov, xi_.x + ri * nij0, xi_.y + ri * nij1);
        } else {
          F0 += -o.t0;
          F1 += -o.t1;
          T += lj0 * -o.t1 - lj1 * -o.t0;
          if (rolling) push_roll(o.m_r);
        }
        if (kAdvance) {
          W.hxi[at] = h;
          W.hroll[at] = hr;
        }
      }
    }
    if (kAdvance) W.cmask[(size_t)lq * N + p] = touch;
    if (A.rec)
      for (int e = 0; e < S; ++e)
        if (e < e0 || e >= cnt || !(touch >> e & 1u)) clear_record(A.rec, (size_t)e * N + p);

    // walls, then hull (engine.py:420-421, 509-520)
    const double4 *wold = W.wst + (size_t)A.spar * ((size_t)N * (P.n_walls + P.n_hull) + 1);
    double4 *wnew = W.wst + (size_t)(A.spar ^ 1) * ((size_t)N * (P.n_walls + P.n_hull) + 1);
    int kw = 0, kh = 0;
    const int nseg = P.n_walls + P.n_hull;
    for (int s = 0; s < nseg; ++s) {
      const bool wall = s < P.n_walls;
      const double *g = segment(P, A, s);
      SegHit hh;
      const size_t rat = (size_t)(S + s) * N + p;
      const bool touching = s < 64 ? (smask >> s & 1ull) != 0ull
                                   : seg_near(W.segbox + 4 * s, xp.x, xp.y, rp) &&
                                         seg_hit(g, xp.x, xp.y, rp, hh);
      if (!touching) {
        if (A.rec) clear_record(A.rec, rat);
        continue;
      }
      if (s < 64) seg_hit(g, xp.x, xp.y, rp, hh);
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
      if (rolling) push_roll(-o.m_r);
      if (kSums) {
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
      }
      if (kAdvance) wnew[rank] = st;
      if (A.rec)
        put_record(A.rec, rat, (double)p, -1.0, wall ? 1.0 : 2.0, n0, n1, hh.cp0 - xp.x,
                   hh.cp1 - xp.y, o, hh.ov, hh.cp0, hh.cp1);
    }

    // rolling torques after every lever torque, in the same contact order
    if (rolling) {
      if (nroll > 0) T += rr0;
      if (nroll > 1) T += rr1;
      if (nroll > 2) T += rr2;
      if (nroll > 3) T += rr3;
      for (int k = 4; k < nroll; ++k) T += W.mr[(size_t)(k - 4) * N + p];
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
      if (bounded) packed_new = boundary_counts(P, A, xs2[0], xs2[1], rp);
    }
  }
  // boundary contacts at the new positions, for the next launch's ranks:
  // a wave with any adds its total to the block's count (cleared by the prep
  // kernel of this launch)
  if (kAdvance && bounded && __ballot(packed_new != 0)) {
    int v = packed_new;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
    if ((threadIdx.x & 63) == 0) atomicAdd(&W.blkcnt[(size_t)(A.spar ^ 1) * P.nblocks + lb], v);
  }

  // ---- ledger partials, reduced by the next launch (or the flush) --------
  if (kSums) {
    block_sum<kDch>(ch, red);
    if (threadIdx.x == 0)
      for (int k = 0; k < kDch; ++k) W.part[(size_t)k * gridDim.x + lb] = ch[k];
  }
}

__global__ void disk_segbox_kernel(tdm_disk_params P, DiskArgs A)
{
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= P.n_walls + P.n_hull) return;
  const double *g = segment(P, A, s);
  double *b = A.w.segbox + 4 * s;
  b[0] = fmin(g[0], g[2]);
  b[1] = fmin(g[1], g[3]);
  b[2] = fmax(g[0], g[2]);
  b[3] = fmax(g[1], g[3]);
}

__global__ void disk_init_kernel(DiskDyn *dyn)
{
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    DiskDyn d = {};
    d.flag[0] = 1;
    *dyn = d;
  }
}

__global__ void disk_mark_kernel(DiskDyn *dyn, int par)
{
  if (threadIdx.x == 0 && blockIdx.x == 0) dyn->flag[par] = 1;
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

// Kernel arguments of launch i of a call: state buffers alternate from x/v/
// omega (i even), parities continue the caller's launch and step counts.
DiskArgs disk_args(const tdm_disk_arrays *a, const DiskWs &w, int i, int mode, bool records)
{
  DiskArgs A;
  const bool odd = i & 1;
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
  A.par = (int)((a->launches + i) & 1);
  A.spar = (int)((a->steps + (mode & 2 ? i : 0)) & 1);
  A.mode = mode;
  return A;
}

int launch_disk_steps(const tdm_disk_arrays *a, const tdm_disk_params *p, const DiskWs &w,
                      int nsteps, int flags, hipStream_t s, hipEvent_t *ev = nullptr)
{
  const int nb = p->nblocks;
  const int mode = 1 | 2 | (flags & TDM_DISK_LEDGER ? 4 : 0);
  if (nsteps > 0 && p->n_walls + p->n_hull > 0)
    disk_count_kernel<<<nb, kThreads, 0, s>>>(*p, disk_args(a, w, 0, mode, false));
  for (int i = 0; i < nsteps; ++i) {
    const bool rec = (flags & TDM_DISK_RECORD) && a->records && i == nsteps - 1;
    const DiskArgs A = disk_args(a, w, i, mode, rec);
    if (ev) HIP_TRY(hipEventRecord(ev[4 * i], s));
    disk_prep_kernel<<<min(nb, kPrepBlocks), kThreads, 0, s>>>(*p, A);
    if (ev) HIP_TRY(hipEventRecord(ev[4 * i + 1], s));
    if (ev) HIP_TRY(hipEventRecord(ev[4 * i + 2], s));
    if ((flags & TDM_DISK_LEDGER) || i == nsteps - 1)
      disk_step_kernel<true, true><<<nb, kThreads, 0, s>>>(*p, A);
    else
      disk_step_kernel<true, false><<<nb, kThreads, 0, s>>>(*p, A);
    if (ev) HIP_TRY(hipEventRecord(ev[4 * i + 3], s));
  }
  if (nsteps > 0) disk_flush_kernel<<<1, kThreads, 0, s>>>(*p, disk_args(a, w, 0, mode, false));
  HIP_TRY(hipGetLastError());
  return TDM_OK;
}

// Bounding boxes of the wall and hull segments into the workspace, and the
// first kParamSegs of them into the launch parameters (they travel with every
// launch); waits for the copy.
int segment_boxes(const tdm_disk_arrays *a, tdm_disk_params *p, const DiskWs &w,
                  hipStream_t s)
{
  const int nseg = p->n_walls + p->n_hull;
  if (nseg > 0)
    disk_segbox_kernel<<<(nseg + 63) / 64, 64, 0, s>>>(*p, disk_args(a, w, 0, 0, false));
  HIP_TRY(hipGetLastError());
  for (int k = 0; k < kParamSegs; ++k)
    for (int c = 0; c < 4; ++c) p->seg_box[k][c] = 0.0;
  if (nseg > 0) {
    HIP_TRY(hipMemcpyAsync(p->seg_box, w.segbox, sizeof(double) * 4 * min(nseg, kParamSegs),
                           hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
  }
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

}  // extern "C"

extern "C" {

int tdm_disk_blocks(int n) { return n <= 0 ? 0 : (n + kThreads - 1) / kThreads; }

size_t tdm_disk_workspace_bytes(int n, int ncells, int cell_cap, int cand_slots, int n_walls,
                                int n_hull)
{
  DiskWs w;
  size_t total = 0;
  carve(nullptr, n, ncells, cell_cap, cand_slots, n_walls, n_hull, &w, &total);
  return total;
}

int tdm_disk_setup(tdm_disk_arrays *a, tdm_disk_params *p, void *stream)
{
  DiskWs w;
  int rc = check_disk(a, p, &w);
  if (rc) return rc;
  hipStream_t s = static_cast<hipStream_t>(stream);
  const size_t ncells = (size_t)p->ncell[0] * p->ncell[1];
  fill_kernel<int><<<256, 256, 0, s>>>(w.cnt, 2 * (size_t)p->n, 0);
  fill_kernel<unsigned><<<256, 256, 0, s>>>(w.cmask, 2 * (size_t)p->n, 0u);
  fill_kernel<int><<<256, 256, 0, s>>>(w.ccount, 2 * ncells, 0);
  disk_init_kernel<<<1, 64, 0, s>>>(w.dyn);
  HIP_TRY(hipGetLastError());
  rc = segment_boxes(a, p, w, s);
  if (rc) return rc;
  a->launches = 0;
  a->steps = 0;
  return TDM_OK;
}

int tdm_disk_segments(tdm_disk_arrays *a, tdm_disk_params *p, void *stream)
{
  DiskWs w;
  int rc = check_disk(a, p, &w);
  if (rc) return rc;
  return segment_boxes(a, p, w, static_cast<hipStream_t>(stream));
}

int tdm_disk_rebin(tdm_disk_arrays *a, const tdm_disk_params *p, void *stream)
{
  DiskWs w;
  int rc = check_disk(a, p, &w);
  if (rc) return rc;
  disk_mark_kernel<<<1, 64, 0, static_cast<hipStream_t>(stream)>>>(w.dyn,
                                                                    (int)(a->launches & 1));
  HIP_TRY(hipGetLastError());
  return TDM_OK;
}

int tdm_disk_step(tdm_disk_arrays *a, const tdm_disk_params *p, int nsteps, int flags,
                  void *stream)
{
  DiskWs w;
  int rc = check_disk(a, p, &w);
  if (rc) return rc;
  if (nsteps < 0) return fail(TDM_EINVAL, "nsteps must be nonnegative");
  rc = launch_disk_steps(a, p, w, nsteps, flags, static_cast<hipStream_t>(stream));
  if (rc == TDM_OK) {
    a->launches += nsteps;
    a->steps += nsteps;
  }
  return rc;
}

int tdm_disk_inspect(tdm_disk_arrays *a, const tdm_disk_params *p, void *stream)
{
  DiskWs w;
  int rc = check_disk(a, p, &w);
  if (rc) return rc;
  hipStream_t s = static_cast<hipStream_t>(stream);
  const DiskArgs A = disk_args(a, w, 0, 1, a->records != nullptr);
  if (p->n_walls + p->n_hull > 0) disk_count_kernel<<<p->nblocks, kThreads, 0, s>>>(*p, A);
  disk_prep_kernel<<<min(p->nblocks, kPrepBlocks), kThreads, 0, s>>>(*p, A);
  disk_step_kernel<false, true><<<p->nblocks, kThreads, 0, s>>>(*p, A);
  disk_flush_kernel<<<1, kThreads, 0, s>>>(*p, A);
  HIP_TRY(hipGetLastError());
  a->launches += 1;
  return TDM_OK;
}

int tdm_disk_step_timed(tdm_disk_arrays *a, const tdm_disk_params *p, int nsteps,
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
  if (rc == TDM_OK) {
    a->launches += nsteps;
    a->steps += nsteps;
  }
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
