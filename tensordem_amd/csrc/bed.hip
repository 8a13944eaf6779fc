// One step of unbonded Hertz-Mindlin disk beds on MI355X (gfx950).
//
// The contact law restates, in two dimensions, the reference's floe-bed
// kernel (src/tensordem/scale_ceiling/kernel.py:277-401): Hertz normal force
// (2/3) s_n d with s_n = 2 E* sqrt(R* d), viscous term
// 2 sqrt(5/6) beta sqrt(s_n m*) v_n, non-cohesive clamp; Cundall-Strack
// tangential spring xi += v_t dt with stiffness 8 G* sqrt(R* d) scaled back
// to the Coulomb cap; floor and back-wall contacts; gravity.  A pair's spring
// history is dropped once it separates (granular2d/engine.py:388-396).
//
// Layout: every disk owns a sorted list of candidate partners (all disks
// within r_p + r_q + skin when the list was built) and one spring history per
// list slot.  Both disks of a pair keep a copy of the pair's history and
// update it with identical arithmetic (lo -> hi orientation), so the copies
// stay bit-identical and no disk ever writes another disk's data: forces are
// gathered, not scattered, and no float atomics are used.  Lists and
// histories are double-buffered by list parity; a rebuild writes the new
// lists into the spare buffer and carries every active history over by
// merging the old and new partner lists (sorted-key remap).
//
// Per disk the force is gravity, then the pair forces by ascending partner
// index, then floor, then back wall -- the order of the FP64 oracle
// (oracle/bed2d.c), so FP64 runs are bit-identical to it.  A disk's
// candidate indices are prefetched into registers and its partners' state is
// loaded kChunk at a time, so one wave waits on one gather per chunk rather
// than one per partner.
#include "common.h"

#include <vector>

namespace {

using namespace tdm;

constexpr int kMaxSlots = 16;
constexpr int kChunk = 2;        // partners gathered per round
// A disk that moved this fraction of the skin since its list was built
// requests a rebuild; the list stays complete up to half the skin, and the
// bin and search steps follow within two steps (motion per step is far below
// the remaining 0.1 skin; a list used beyond half the skin raises the stale
// status).
constexpr double kRebuildFrac = 0.4;
constexpr int kPre = 6;          // candidate slots fetched before the list length is known
constexpr int kNch = 12;         // ledger channels written per block

enum { MODE_IDLE = 0, MODE_SEARCH = 1, MODE_BIN = 2 };
enum { OP_STEP = 1, OP_ENERGY = 2 };
constexpr int kLpBit = 1 << 16;  // list parity, mirrored into each block-table entry
// ledger partial channels
// C_COUNTS carries pair contacts + 2^32 x wall contacts: both are integers
// far below 2^21, so the sums over blocks stay exact in FP64 (a ninth step
// channel would cost the step kernel a wave per SIMD)
enum {
  C_DASH = 0, C_FRIC, C_WDASH, C_WFRIC, C_WFX, C_WFY, C_COUNTS, C_WWORK,
  C_KIN = 8, C_POT, C_ELN, C_ELT
};
constexpr int kStepCh = 8;       // step channels C_DASH .. C_WWORK
constexpr double kWallCount = 4294967296.0;  // 2^32
constexpr uint32_t kFloorBit = 1u << 16, kBackBit = 1u << 17, kRightBit = 1u << 18;
constexpr uint32_t kWallBits = kFloorBit | kBackBit | kRightBit;

// Clock, search phase and list state of a member at the start of a step,
// carried from launch to launch inside one multi-step call (double-buffered
// by the launch parity); the member record is brought up to date when the
// call ends.
struct BedDyn {
  double time;
  long long step;
  int mode, tab, lp, rebuilds, fresh, pad;
};

struct BedWs {
  int4 *blocks;      // (member, first disk, count | list parity << 16, rank)
  double *partials;  // [2][kNch][nblocks], by step parity, channel-major
  int *bflags;       // [nblocks]
  int *cell_count;   // [2][ncells]
  int *cell_tab;     // [2][ncells][cap]
  int *cand;         // [2][slots][n]
  double2 *xi;       // [2][slots][n]
  uint32_t *mask;    // [2][n]
  uint8_t *ncand;    // [2][n]
  double2 *xb;       // [n] position when the list was built
  BedDyn *dyn;       // [nmembers][2]
  int *mflags;       // [nmembers][4]: status | request << 16 of steps s % 3
};

// Spring histories are stored in the precision of the forces: FP64 steps keep
// double2, FP32 steps float2 (their histories are computed in FP32, so the
// narrower store is exact) in the first half of the same region.
template <typename R> struct HistT { using type = double2; };
template <> struct HistT<float> { using type = float2; };
template <typename R> using Hist = typename HistT<R>::type;
__device__ inline double2 widen(double2 h) { return h; }
__device__ inline double2 widen(float2 h) { return make_double2((double)h.x, (double)h.y); }
template <typename H> __device__ inline H narrow(double2 h);
template <> __device__ inline double2 narrow<double2>(double2 h) { return h; }
template <> __device__ inline float2 narrow<float2>(double2 h)
{
  return make_float2((float)h.x, (float)h.y);
}

// Carves the workspace at `base`; with base == nullptr only the byte offsets
// are wanted and the returned pointers hold them (used by the layout query).
inline void carve(void *base, int n, int nblocks, int nmembers, int ncells, int slots, int cap,
                  BedWs *w, size_t *total)
{
  size_t off = 0;
  const uintptr_t b = reinterpret_cast<uintptr_t>(base);
  auto take = [&](size_t bytes) {
    char *p = reinterpret_cast<char *>(b + off);
    off += align256(bytes);
    return p;
  };
  w->blocks = reinterpret_cast<int4 *>(take(sizeof(int4) * size_t(nblocks)));
  w->partials = reinterpret_cast<double *>(take(sizeof(double) * 2 * kNch * size_t(nblocks)));
  w->bflags = reinterpret_cast<int *>(take(sizeof(int) * size_t(nblocks)));
  w->cell_count = reinterpret_cast<int *>(take(sizeof(int) * 2 * size_t(ncells)));
  w->cell_tab = reinterpret_cast<int *>(take(sizeof(int) * 2 * size_t(ncells) * size_t(cap)));
  w->cand = reinterpret_cast<int *>(take(sizeof(int) * 2 * size_t(slots) * size_t(n)));
  w->xi = reinterpret_cast<double2 *>(take(sizeof(double2) * 2 * size_t(slots) * size_t(n)));
  w->mask = reinterpret_cast<uint32_t *>(take(sizeof(uint32_t) * 2 * size_t(n)));
  w->ncand = reinterpret_cast<uint8_t *>(take(2 * size_t(n)));
  w->xb = reinterpret_cast<double2 *>(take(sizeof(double2) * size_t(n)));
  w->dyn = reinterpret_cast<BedDyn *>(take(sizeof(BedDyn) * 2 * size_t(nmembers)));
  w->mflags = reinterpret_cast<int *>(take(sizeof(int) * 4 * size_t(nmembers)));
  *total = off;
}

struct BKArgs {
  tdm_bed_arrays a;
  BedWs w;
  const double2 *x_in, *v_in;
  double2 *x_out, *v_out;
  int op;
  int parity;  // 0: x, v -> x_alt, v_alt; 1: the reverse
  int first;   // first launch of a multi-step call: the member record holds the state
};

// The member constants one step reads, moved to scalar registers (every lane
// of a workgroup belongs to the same member).
struct BedConst {
  double e_star, g_star, damping, friction, wall_friction, rho_pi_t, gravity_y, dt;
  double floor_y, back_wall_x, right_x, right_v, time, skin, cell_size, ox, oy;
  int use_floor, use_back_wall, use_right, ncx, ncy, cell_first, block_first, nblocks, mode, tab;
};

__device__ inline double uniform(double v)
{
  return __hiloint2double(__builtin_amdgcn_readfirstlane(__double2hiint(v)),
                          __builtin_amdgcn_readfirstlane(__double2loint(v)));
}
__device__ inline int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ inline long long uniform(long long v)
{
  const uint32_t hi = __builtin_amdgcn_readfirstlane((int)(v >> 32));
  const uint32_t lo = __builtin_amdgcn_readfirstlane((int)v);
  return (long long)(((uint64_t)hi << 32) | lo);
}

__device__ inline BedConst load_const(const tdm_bed_member *mp)
{
  BedConst c;
  c.e_star = uniform(mp->e_star);
  c.g_star = uniform(mp->g_star);
  c.damping = uniform(mp->damping);
  c.friction = uniform(mp->friction);
  c.wall_friction = uniform(mp->wall_friction);
  c.rho_pi_t = uniform(mp->rho_pi_t);
  c.gravity_y = uniform(mp->gravity_y);
  c.dt = uniform(mp->dt);
  c.floor_y = uniform(mp->floor_y);
  c.back_wall_x = uniform(mp->back_wall_x);
  c.right_x = uniform(mp->right_wall_x);
  c.right_v = uniform(mp->right_wall_v);
  c.time = uniform(mp->time);
  c.skin = uniform(mp->skin);
  c.cell_size = uniform(mp->cell_size);
  c.ox = uniform(mp->grid_origin[0]);
  c.oy = uniform(mp->grid_origin[1]);
  c.use_floor = uniform(mp->use_floor);
  c.use_back_wall = uniform(mp->use_back_wall);
  c.use_right = uniform(mp->use_right_wall);
  c.ncx = uniform(mp->ncx);
  c.ncy = uniform(mp->ncy);
  c.cell_first = uniform(mp->cell_first);
  c.block_first = uniform(mp->block_first);
  c.nblocks = uniform(mp->nblocks);
  c.mode = uniform(mp->mode);
  c.tab = uniform(mp->tab);
  return c;
}

// The pair-law constants of a member, in scalar registers for the whole step
// (the rest of the member record is read where it is used, so the step
// kernel's scalar register budget holds without spills).
struct BedLaw {
  double e_star, g_star, damping, friction, rho_pi_t, dt;
};

__device__ inline BedLaw load_law(const tdm_bed_member *mp)
{
  BedLaw c;
  c.e_star = uniform(mp->e_star);
  c.g_star = uniform(mp->g_star);
  c.damping = uniform(mp->damping);
  c.friction = uniform(mp->friction);
  c.rho_pi_t = uniform(mp->rho_pi_t);
  c.dt = uniform(mp->dt);
  return c;
}

// FP32 copies of the law constants, formed once per thread.
struct BedLaw32 {
  float two_e, eight_g, inv_eight_g, damping, friction, rho_pi_t, dt;
};

__device__ inline BedLaw32 law32(const BedLaw &m)
{
  const float eight_g = 8.0f * (float)m.g_star;
  return BedLaw32{2.0f * (float)m.e_star, eight_g, 1.0f / eight_g, (float)m.damping,
                  (float)m.friction, (float)m.rho_pi_t, (float)m.dt};
}

// Pair law of one overlapping pair, lo -> hi oriented: (dx, dy) = x_hi - x_lo
// and s2 = dx^2 + dy^2 as the contact test formed them, rs = r_p + r_q,
// rs2 = rs * rs, (vrx, vry) = v_lo - v_hi.  The radii enter only through
// r_p r_q, r_p + r_q and the masses' product and sum, so either disk of a pair
// may pass them in its own order and get the same bits.  On return F is the
// force on hi (lo receives -F) and xi the pair's updated spring history.
//
// FP64: the reference's arithmetic (oracle/bed2d.c), bit for bit.  FP32: the
// overlap d = rs - |dx| is formed as (rs^2 - s2) / (rs + |dx|), with the
// numerator in FP64, so no FP64 square root is needed (relative error a few
// FP32 ulps), and roots and reciprocals share hardware rsq/rcp results;
// whether the pair touches was decided in FP64 beforehand.
__device__ inline void hm_pair64(const BedLaw &m, double dx, double dy, double dist, double rs,
                                 double rp, double rq, double vrx, double vry, double2 &xi,
                                 double &Fx, double &Fy, double &dash, double &fric)
{
  const double d = rs - dist;
  const double dd = dist > 1e-300 ? dist : 1e-300;
  const double nx = dx / dd, ny = dy / dd;
  const double reff = (rp * rq) / rs;
  const double mlo = m.rho_pi_t * (rp * rp), mhi = m.rho_pi_t * (rq * rq);
  const double meff = (mlo * mhi) / (mlo + mhi);
  const double sn = (2 * m.e_star) * sqrt(reff * d);
  const double fel = ((2.0 / 3.0) * sn) * d;
  const double vn = vrx * nx + vry * ny;
  double fn = fel + (m.damping * sqrt(sn * meff)) * vn;
  if (fn < 0) fn = 0;
  const double st = (8 * m.g_star) * sqrt(reff * d);
  const double vtx = vrx - vn * nx, vty = vry - vn * ny;
  double ex = xi.x + vtx * m.dt, ey = xi.y + vty * m.dt;
  const double mag = sqrt(ex * ex + ey * ey);
  const double cap = m.friction * fn;
  if (mag > 0) {
    const double den = st * mag;
    double sc = cap / (den > 1e-300 ? den : 1e-300);
    if (sc > 1) sc = 1;
    ex = ex * sc;
    ey = ey * sc;
  }
  xi = make_double2(ex, ey);
  Fx = fn * nx + st * ex;
  Fy = fn * ny + st * ey;
  dash = (fn - fel) * vn;
  fric = st * (ex * vrx + ey * vry);
}

__device__ inline void hm_pair32(const BedLaw32 &m, double dx, double dy, double s2, double rs,
                                 double rs2, float rpf, float rqf, double vrx, double vry,
                                 float2 &xi, double &Fx, double &Fy, double &dash, double &fric)
{
  // six transcendental operations per pair: 1/|dx| and |dx| from one rsq,
  // R* and m* over one common reciprocal, sqrt(R* d) and its inverse from
  // one rsq (the inverse feeds the Coulomb scale), the damping root and the
  // spring's inverse length
  const float s2f = (float)s2;
  const float inv = rsq32(fmaxf(s2f, 1e-37f));
  const float distf = s2f * inv;
  const float df = fmaxf((float)(rs2 - s2) * rcp32((float)rs + distf), 0.0f);
  const float nx = (float)dx * inv, ny = (float)dy * inv;
  const float rp2 = rpf * rpf, rq2 = rqf * rqf;
  const float rsum = rpf + rqf, r2sum = rp2 + rq2;
  const float rden = rcp32(rsum * r2sum);
  const float reff = ((rpf * rqf) * r2sum) * rden;
  const float meff = (m.rho_pi_t * (rp2 * rq2)) * (rsum * rden);
  const float t = reff * df;
  const float it = rsq32(fmaxf(t, 1e-37f));
  const float sq = t * it;
  const float sn = m.two_e * sq;
  const float fel = ((2.0f / 3.0f) * sn) * df;
  const float vxf = (float)vrx, vyf = (float)vry;
  const float vn = vxf * nx + vyf * ny;
  const float fn = fmaxf(fel + (m.damping * sqrt32(sn * meff)) * vn, 0.0f);
  const float st = m.eight_g * sq;
  const float vtx = vxf - vn * nx, vty = vyf - vn * ny;
  float ex = xi.x + vtx * m.dt, ey = xi.y + vty * m.dt;
  const float m2 = ex * ex + ey * ey;
  if (m2 > 0) {
    // |xi| s_t capped at mu f_n: scale = mu f_n / (s_t |xi|)
    const float sc = fminf(((m.friction * fn) * (it * m.inv_eight_g)) * rsq32(m2), 1.0f);
    ex *= sc;
    ey *= sc;
  }
  xi = make_float2(ex, ey);
  Fx = (double)(fn * nx + st * ex);
  Fy = (double)(fn * ny + st * ey);
  dash = (double)((fn - fel) * vn);
  fric = (double)(st * (ex * vxf + ey * vyf));
}

// Wall contact along the inward unit normal (nx, ny) at distance `gap`.
template <typename R>
__device__ inline bool hm_wall(const BedLaw &m, double wall_friction, double r, double mass,
                               double gap, double nx, double ny, double2 v, double2 &xi,
                               double &Fx, double &Fy, double &dash, double &fric)
{
  const double d = r - gap;
  if (!(d > 0)) return false;
  if constexpr (Prec<R>::fp64) {
    const double sn = (2 * m.e_star) * sqrt(r * d);
    const double fel = ((2.0 / 3.0) * sn) * d;
    const double vn = v.x * nx + v.y * ny;
    double fn = fel - (m.damping * sqrt(sn * mass)) * vn;
    if (fn < 0) fn = 0;
    const double st = (8 * m.g_star) * sqrt(r * d);
    const double vtx = v.x - vn * nx, vty = v.y - vn * ny;
    double ex = xi.x + vtx * m.dt, ey = xi.y + vty * m.dt;
    const double mag = sqrt(ex * ex + ey * ey);
    const double cap = wall_friction * fn;
    if (mag > 0) {
      const double den = st * mag;
      double sc = cap / (den > 1e-300 ? den : 1e-300);
      if (sc > 1) sc = 1;
      ex = ex * sc;
      ey = ey * sc;
    }
    xi = make_double2(ex, ey);
    Fx = fn * nx - st * ex;
    Fy = fn * ny - st * ey;
    dash = (fn - fel) * (-vn);
    fric = st * (ex * v.x + ey * v.y);
  } else {
    const float df = (float)d, rf = (float)r;
    const float sq = sqrt32(rf * df);
    const float sn = (2.0f * (float)m.e_star) * sq;
    const float fel = ((2.0f / 3.0f) * sn) * df;
    const float vn = (float)v.x * (float)nx + (float)v.y * (float)ny;
    const float fn = fmaxf(fel - ((float)m.damping * sqrt32(sn * (float)mass)) * vn, 0.0f);
    const float st = (8.0f * (float)m.g_star) * sq;
    const float vtx = (float)v.x - vn * (float)nx, vty = (float)v.y - vn * (float)ny;
    const float dtf = (float)m.dt;
    float ex = (float)xi.x + vtx * dtf, ey = (float)xi.y + vty * dtf;
    const float mag = sqrt32(ex * ex + ey * ey);
    const float cap = (float)wall_friction * fn;
    if (mag > 0) {
      const float sc = fminf(cap * rcp32(fmaxf(st * mag, 1e-30f)), 1.0f);
      ex *= sc;
      ey *= sc;
    }
    xi = make_double2((double)ex, (double)ey);
    Fx = (double)(fn * (float)nx - st * ex);
    Fy = (double)(fn * (float)ny - st * ey);
    dash = (double)((fn - fel) * (-vn));
    fric = (double)(st * (ex * (float)v.x + ey * (float)v.y));
  }
  return true;
}

// Cell of a position, clamped into the member's grid (non-finite -> cell 0).
__device__ inline int cell_of(double ox, double oy, double h, int ncx, int ncy, int first,
                              double2 x, int *cx, int *cy)
{
  int ix = (int)floor((x.x - ox) / h);
  int iy = (int)floor((x.y - oy) / h);
  if (!(x.x == x.x) || !(x.y == x.y)) { ix = 0; iy = 0; }
  ix = min(max(ix, 0), ncx - 1);
  iy = min(max(iy, 0), ncy - 1);
  *cx = ix;
  *cy = iy;
  return first + iy * ncx + ix;
}

// Rebuild of disk p's candidate list (the step after the hash was filled):
// all disks within r_p + r_q + skin in the 3x3 cells around p, ascending,
// written to list buffer lp^1, with the spring history of every pair that
// was in contact carried over by a sorted merge of the old and new lists.
// Both lists stay in registers (static slot indices after unrolling); the
// cell entries are visited kScan at a time, so a wave waits on one table
// load and one partner gather per batch rather than per entry.  On return
// cq holds the new list (slots past the length hold p).  Returns the new
// length and mask; *status gets CAND_OVERFLOW if S is short.
constexpr int kScan = 4;

// The member's contact grid, read when a step bins or searches.
struct BedGrid {
  double ox, oy, h;
  int ncx, ncy, first;
};

__device__ inline BedGrid load_grid(const tdm_bed_member *mp)
{
  BedGrid g;
  g.ox = uniform(mp->grid_origin[0]);
  g.oy = uniform(mp->grid_origin[1]);
  g.h = uniform(mp->cell_size);
  g.ncx = uniform(mp->ncx);
  g.ncy = uniform(mp->ncy);
  g.first = uniform(mp->cell_first);
  return g;
}

template <typename R>
__device__ inline int rebuild_list(const BKArgs &k, const BedGrid &m, double skin, int tb,
                                   int p, double2 xp, double rp, int lp, int nc, uint32_t mask,
                                   int (&cq)[kPre], uint32_t *nmask_out, int *status)
{
  const tdm_bed_arrays &a = k.a;
  const int n = a.n, S = a.cand_slots;
  const size_t sn = (size_t)S * n;
  const int *cand = k.w.cand + lp * sn + p;
  const Hist<R> *xis = reinterpret_cast<const Hist<R> *>(k.w.xi) + lp * sn + p;
  int *ncl = k.w.cand + (lp ^ 1) * sn + p;
  Hist<R> *nxi = reinterpret_cast<Hist<R> *>(k.w.xi) + (lp ^ 1) * sn + p;
  int cx, cy;
  cell_of(m.ox, m.oy, m.h, m.ncx, m.ncy, m.first, xp, &cx, &cy);
  const int *cntab = k.w.cell_count + (size_t)tb * a.ncells;
  const int *tab = k.w.cell_tab + (size_t)tb * a.ncells * a.cell_cap;
  // entries of the 3x3 block, flattened: cells o = 0..8 hold pre[o]..pre[o+1]-1,
  // and entry t of cell o sits at table offset base[o] + t
  int pre[10], base[9];
  pre[0] = 0;
#pragma unroll
  for (int o = 0; o < 9; ++o) {
    const int gx = cx + o % 3 - 1, gy = cy + o / 3 - 1;
    const bool in = gx >= 0 && gx < m.ncx && gy >= 0 && gy < m.ncy;
    const int cell = in ? m.first + gy * m.ncx + gx : 0;
    pre[o + 1] = in ? min(cntab[cell], a.cell_cap) : 0;
    base[o] = cell * a.cell_cap;
  }
#pragma unroll
  for (int o = 0; o < 9; ++o) {
    base[o] -= pre[o];
    pre[o + 1] += pre[o];
  }
  const int total = pre[9];
  int nl[kMaxSlots];
#pragma unroll
  for (int e = 0; e < kMaxSlots; ++e) nl[e] = INT_MAX;
  int nn = 0;
  for (int t0 = 0; t0 < total; t0 += kScan) {
    int qs[kScan];
#pragma unroll
    for (int u = 0; u < kScan; ++u) {
      const int t = t0 + u;
      qs[u] = p;
      if (t < total) {
        // the cell holding entry t is the last one starting at or before t
        // (an empty cell starts where the next one does, which overrides it)
        int off = base[0];
#pragma unroll
        for (int c = 1; c < 9; ++c) off = t >= pre[c] ? base[c] : off;
        qs[u] = tab[(size_t)(off + t)];
      }
    }
    double2 xq[kScan];
    double rq[kScan];
#pragma unroll
    for (int u = 0; u < kScan; ++u) {
      xq[u] = xp;
      rq[u] = 0.0;
      if (qs[u] != p) { xq[u] = k.x_in[qs[u]]; rq[u] = a.radius[qs[u]]; }
    }
#pragma unroll
    for (int u = 0; u < kScan; ++u) {
      const int q = qs[u];
      if (q == p) continue;
      const double dx = xq[u].x - xp.x, dy = xq[u].y - xp.y;
      // candidates: within r_p + r_q + skin (the slightly inclusive squared
      // form only ever adds pairs, which the contact test then rejects)
      const double cut = (rp + rq[u]) + skin;
      if (!(dx * dx + dy * dy < (cut * cut) * (1.0 + 1e-15))) continue;
      if (nn == S) { *status |= TDM_ST_CAND_OVERFLOW; continue; }
      // sorted insert: one compare-exchange pass over the register list
      int carry = q;
#pragma unroll
      for (int e = 0; e < kMaxSlots; ++e) {
        const int lo = min(nl[e], carry), hi = max(nl[e], carry);
        nl[e] = lo;
        carry = hi;
      }
      ++nn;
    }
  }
  // carry each old slot that was in contact to its place in the new list
  uint32_t nmask = mask & kWallBits;
#pragma unroll
  for (int o = 0; o < kMaxSlots; ++o) {
    if (o >= nc || !((mask >> o) & 1u)) continue;
    const int q = o < kPre ? cq[o] : cand[(size_t)o * n];
    int pos = 0;
    bool hit = false;
#pragma unroll
    for (int e = 0; e < kMaxSlots; ++e) {
      pos += nl[e] < q;
      hit |= nl[e] == q;
    }
    if (hit) {
      nxi[(size_t)pos * n] = xis[(size_t)o * n];
      nmask |= 1u << pos;
    }
  }
#pragma unroll
  for (int e = 0; e < kMaxSlots; ++e) {
    if (e < nn) ncl[(size_t)e * n] = nl[e];
    if (e < kPre) cq[e] = e < nn ? nl[e] : p;
  }
  k.w.ncand[(size_t)(lp ^ 1) * n + p] = (uint8_t)nn;
  *nmask_out = nmask;
  return nn;
}

template <int U, int G>
__device__ void bed_member_ledger(const BKArgs &k, int mi, long long step, int flags,
                                  double (*wsum)[kCh]);
__device__ inline BedDyn bed_advance(const BedDyn &d, int flags, double dt);

// kDeferred: a launch inside a multi-step call.  After the call's first
// launch the member's clock and search phase come from the previous
// launch's state and flags instead of the member record, each block flips
// its own block-table list parity after a rebuild, and the member's first
// block reduces the previous step's ledger after its own work (one launch
// per step, no finalize launch).
// kReport: the velocity readout: the forces of the current state on the
// current lists, and only the full-step velocity stored (into v_out).
// kVerlet: the batch has velocity-Verlet members (the update is compiled in
// only then: it costs the Euler-only kernels 1 % in registers).
template <typename R, bool kDeferred, bool kReport, bool kVerlet>
__device__ __forceinline__ void bed_step_body(const BKArgs &k)
{
  __shared__ double red[4][kCh];
  __shared__ int sflags[2];
  const tdm_bed_arrays &a = k.a;
  const int4 blk = k.w.blocks[blockIdx.x];
  const int mi = blk.x, p0 = blk.y, cnt = blk.z & 0xFFFF, lp0 = (blk.z & kLpBit) ? 1 : 0;
  const bool valid = (int)threadIdx.x < cnt;
  const int p = p0 + (valid ? (int)threadIdx.x : 0);
  const int n = a.n, S = a.cand_slots;
  const size_t sn = (size_t)S * n;
  const tdm_bed_member *mp = a.members + mi;
  // own state, list header and the first kPre candidate slots in one round
  const double2 xp = k.x_in[p], vp = k.v_in[p];
  const double rp = a.radius[p];
  double2 xb = k.w.xb[p];
  uint32_t mask = k.w.mask[(size_t)lp0 * n + p];
  int nc = k.w.ncand[(size_t)lp0 * n + p];
  int cq[kPre];  // the first list slots; later slots are read when reached
#pragma unroll
  for (int e = 0; e < kPre; ++e) cq[e] = e < S ? k.w.cand[lp0 * sn + (size_t)e * n + p] : p;
  const BedLaw law = load_law(mp);
  const double gravity_y = uniform(mp->gravity_y), skin = uniform(mp->skin);
  int mode = uniform(mp->mode), tab = uniform(mp->tab);
  double time = uniform(mp->time);
  BedDyn pd, cur;
  int pflags = 0;
  if (kDeferred && k.first) {
    // the call's first step starts from the member record
    pd.time = time;
    pd.step = uniform((long long)mp->step_count);
    pd.mode = mode;
    pd.tab = tab;
    pd.lp = uniform(mp->list_parity);
    pd.rebuilds = uniform(mp->rebuilds);
    pd.fresh = 1;
    pd.pad = 0;
    cur = pd;
  } else if (kDeferred) {
    // the state at the start of the previous step and the flags that step
    // raised (written by the previous launch); uniform, scalar loads
    pd = k.w.dyn[2 * mi + (k.parity ^ 1)];
    const int4 pf4 = reinterpret_cast<const int4 *>(k.w.mflags)[mi];
    const int r = (int)(pd.step % 3);
    pflags = r == 0 ? pf4.x : (r == 1 ? pf4.y : pf4.z);
    cur = pd.fresh ? pd : bed_advance(pd, pflags, law.dt);
    mode = cur.mode;
    tab = cur.tab;
    time = cur.time;
  }
  if (kReport) mode = MODE_IDLE;  // the current lists hold every contact
  if (threadIdx.x < 2) sflags[threadIdx.x] = 0;
  // ledger channels, gathered into ch[] after the forces (the pair loop
  // carries only its own two)
  double c_dash = 0.0, c_fric = 0.0, c_wdash = 0.0, c_wfric = 0.0, c_wfx = 0.0, c_wfy = 0.0;
  double c_counts = 0.0, c_wwork = 0.0;
  int status = 0, req = 0;
  int lp = lp0;

  if (valid) {
    const double mass = law.rho_pi_t * (rp * rp);
    if (mode == MODE_SEARCH) {
      uint32_t nmask;
      nc = rebuild_list<R>(k, load_grid(mp), skin, tab, p, xp, rp, lp0, nc, mask, cq, &nmask,
                           &status);
      mask = nmask;
      lp = lp0 ^ 1;
      xb = xp;
      k.w.xb[p] = xp;
    } else {
      const double ddx = xp.x - xb.x, ddy = xp.y - xb.y;
      const double lim = 0.5 * skin * 0.999;
      if (ddx * ddx + ddy * ddy >= lim * lim) status |= TDM_ST_STALE_LIST;
    }
    Hist<R> *xis = reinterpret_cast<Hist<R> *>(k.w.xi) + lp * sn + p;
    const BedLaw32 law_f = law32(law);
    const float rpf = (float)rp;

    // ---- forces: gravity, pairs by ascending partner, floor, back wall.
    // Slots past the list length read the disk itself (cache hits) so that
    // every gather of a chunk is issued unconditionally.
    double fx = 0.0, fy = mass * gravity_y;
    uint32_t nmask = 0;
    int npair = 0, nwall = 0;
#pragma unroll
    for (int c0 = 0; c0 < kMaxSlots; c0 += kChunk) {
      if (!__any(c0 < nc)) break;
      int qq[kChunk];
      double2 xq[kChunk], vq[kChunk];
      double rq[kChunk];
      Hist<R> hq[kChunk];
#pragma unroll
      for (int j = 0; j < kChunk; ++j) {
        const int e = c0 + j;
        if (e < kPre) qq[j] = e < nc ? cq[e] : p;
        else qq[j] = e < nc ? k.w.cand[lp * sn + (size_t)e * n + p] : p;
        xq[j] = k.x_in[qq[j]];
        rq[j] = a.radius[qq[j]];
        // a partner's velocity and the pair's history are needed only in
        // contact: fetched with the round for the pairs in contact last
        // step, on demand for new ones
        if ((mask >> e) & 1u) {
          vq[j] = k.v_in[qq[j]];
          hq[j] = xis[(size_t)e * n];
        }
      }
#pragma unroll
      for (int j = 0; j < kChunk; ++j) {
        const int e = c0 + j;
        if (e >= nc) continue;
        // partner minus self; the lo -> hi difference is this or its
        // negation (exact), and s2 is the same either way
        double dx = xq[j].x - xp.x, dy = xq[j].y - xp.y;
        const double s2 = dx * dx + dy * dy;
        const double rs = rp + rq[j];
        const double rs2 = rs * rs;
        // contact iff rs - sqrt(s2) > 0, decided in FP64 for both precisions:
        // above rs2 (1 + 1e-15) the correctly rounded root is >= rs, below
        // rs2 (1 - 1e-15) it is < rs; only the band between needs the root
        if (s2 > rs2 * (1.0 + 1e-15)) continue;
        double dist = 0.0;
        if constexpr (Prec<R>::fp64) {
          dist = sqrt(s2);
          if (!(rs - dist > 0)) continue;
        } else {
          if (!(s2 < rs2 * (1.0 - 1e-15)) && !(rs - sqrt(s2) > 0)) continue;
        }
        const bool self_lo = qq[j] > p;
        if (!((mask >> e) & 1u)) {
          vq[j] = k.v_in[qq[j]];
          hq[j] = Hist<R>{0, 0};
        }
        double vrx = vp.x - vq[j].x, vry = vp.y - vq[j].y;  // self minus partner
        if (!self_lo) { dx = -dx; dy = -dy; vrx = -vrx; vry = -vry; }
        double Fx, Fy, dash, fric;
        if constexpr (Prec<R>::fp64)
          hm_pair64(law, dx, dy, dist, rs, rp, rq[j], vrx, vry, hq[j], Fx, Fy, dash, fric);
        else
          hm_pair32(law_f, dx, dy, s2, rs, rs2, rpf, (float)rq[j], vrx, vry, hq[j], Fx, Fy,
                    dash, fric);
        if constexpr (!kReport) xis[(size_t)e * n] = hq[j];
        nmask |= 1u << e;
        if (self_lo) {
          fx += -Fx; fy += -Fy;
          c_dash += dash; c_fric += fric; ++npair;
        } else {
          fx += Fx; fy += Fy;
        }
      }
    }
    // floor, back wall, right wall: one loop so the law is inlined once
    double2 *xw = reinterpret_cast<double2 *>(a.xi_wall) + 3 * p;
    const int use_floor = uniform(mp->use_floor), use_back = uniform(mp->use_back_wall);
    const int use_right = uniform(mp->use_right_wall);
    const double floor_y = uniform(mp->floor_y), back_x = uniform(mp->back_wall_x);
    const double right_x = uniform(mp->right_wall_x), right_v = uniform(mp->right_wall_v);
    const double wall_mu = uniform(mp->wall_friction);
#pragma unroll 1
    for (int w = 0; w < 3; ++w) {
      if (!(w == 0 ? use_floor : (w == 1 ? use_back : use_right))) continue;
      // the law sees the disk's velocity relative to the wall (kernel.py:381-389);
      // the floor and the back wall stand still
      const double wv = w == 2 ? right_v : 0.0;
      const double gap = w == 0 ? xp.y - floor_y
                                : (w == 1 ? xp.x - back_x : (right_x + right_v * time) - xp.x);
      const double nx = w == 0 ? 0.0 : (w == 1 ? 1.0 : -1.0), ny = w == 0 ? 1.0 : 0.0;
      const uint32_t bit = kFloorBit << w;
      double2 h = (mask & bit) ? xw[w] : make_double2(0.0, 0.0);
      const double2 vr = make_double2(vp.x - wv, vp.y);
      double Fx, Fy, dash, fric;
      if (hm_wall<R>(law, wall_mu, rp, mass, gap, nx, ny, vr, h, Fx, Fy, dash, fric)) {
        if constexpr (!kReport) xw[w] = h;
        nmask |= bit;
        fx += Fx; fy += Fy;
        c_wdash += dash; c_wfric += fric; c_wfx += Fx; c_wfy += Fy;
        c_wwork += Fx * wv;
        ++nwall;
      }
    }
    if constexpr (kReport) {
      // the full-step velocity: v itself, or v + (dt/2 F) / m for a
      // velocity-Verlet state (the reference's second half kick)
      double2 vf = vp;
      if (uniform(mp->integrator) && !uniform(mp->verlet_start)) {
        const double hk = 0.5 * law.dt, inv_m = 1.0 / mass;
        vf = make_double2(vp.x + (hk * fx) * inv_m, vp.y + (hk * fy) * inv_m);
      }
      k.v_out[p] = vf;
    } else {
      // the mask changes only when a contact forms or breaks (a search step
      // writes the new list's buffer): most lanes skip the store
      if (mode == MODE_SEARCH || nmask != mask) k.w.mask[(size_t)lp * n + p] = nmask;
      c_counts = (double)npair + kWallCount * (double)nwall;

      // velocity-Verlet members: v is a full-step velocity only on a new or
      // edited state, which only a call's first step sees (read here, where
      // the scalar registers are free)
      double vx, vy;
      if (kVerlet && uniform(mp->integrator)) {
        const bool vfull = (!kDeferred || k.first) && uniform(mp->verlet_start) != 0;
        // ---- velocity-Verlet (scale_ceiling/kernel.py:410-421): the second
        // half kick of the previous step (none from a full-step velocity), the
        // first half kick of this one, each (dt/2 F) / m in the reference's
        // order, then the drift below
        const double hk = 0.5 * law.dt, inv_m = 1.0 / mass;
        double ux = vp.x, uy = vp.y;
        if (!vfull) {
          ux = vp.x + (hk * fx) * inv_m;
          uy = vp.y + (hk * fy) * inv_m;
        }
        vx = ux + (hk * fx) * inv_m;
        vy = uy + (hk * fy) * inv_m;
      } else if constexpr (Prec<R>::fp64) {
        // ---- semi-implicit Euler (dem.py:363-365) -------------------------
        vx = vp.x + (law.dt * fx) / mass;
        vy = vp.y + (law.dt * fy) / mass;
      } else {
        const double dt_m = law.dt / mass;
        vx = vp.x + fx * dt_m;
        vy = vp.y + fy * dt_m;
      }
      const double xn = xp.x + law.dt * vx, yn = xp.y + law.dt * vy;
      k.x_out[p] = make_double2(xn, yn);
      k.v_out[p] = make_double2(vx, vy);
      if (!isfinite(xn) || !isfinite(yn) || !isfinite(vx) || !isfinite(vy)) status |= TDM_ST_NONFINITE;
      const double ddx = xn - xb.x, ddy = yn - xb.y;
      const double q4 = kRebuildFrac * skin;
      if (ddx * ddx + ddy * ddy >= q4 * q4) req = 1;
      if (mode == MODE_BIN) {
        const BedGrid g = load_grid(mp);
        int cx, cy;
        const int cell = cell_of(g.ox, g.oy, g.h, g.ncx, g.ncy, g.first, make_double2(xn, yn), &cx,
                                 &cy);
        const int t = tab ^ 1;
        const int slot = atomicAdd(k.w.cell_count + (size_t)t * a.ncells + cell, 1);
        if (slot < a.cell_cap) k.w.cell_tab[((size_t)t * a.ncells + cell) * a.cell_cap + slot] = p;
        else status |= TDM_ST_CELL_OVERFLOW;
      }
    }
  }
  if constexpr (kReport) return;
  if (mode == MODE_SEARCH) {
    // clear the other hash table for the next bin phase
    const int ncell_m = uniform(mp->ncx) * uniform(mp->ncy);
    const int nblocks_m = uniform(mp->nblocks);
    int *cntc = k.w.cell_count + (size_t)(tab ^ 1) * a.ncells + uniform(mp->cell_first);
    for (int c = blk.w * kThreads + threadIdx.x; c < ncell_m; c += nblocks_m * kThreads)
      cntc[c] = 0;
  }
  __syncthreads();
  if (status) atomicOr(&sflags[0], status);
  if (req) atomicOr(&sflags[1], 1);
  double ch[kStepCh];
  ch[C_DASH] = c_dash; ch[C_FRIC] = c_fric; ch[C_WDASH] = c_wdash; ch[C_WFRIC] = c_wfric;
  ch[C_WFX] = c_wfx; ch[C_WFY] = c_wfy; ch[C_COUNTS] = c_counts; ch[C_WWORK] = c_wwork;
  block_sum(ch, red);
  const int pslot = kDeferred ? (int)(cur.step & 1) : 0;
  if (threadIdx.x == 0) {
    // partials by the member's logical block index: the reduction order does
    // not depend on where the member sits in the launch
    const size_t slot = (size_t)(uniform(mp->block_first) + blk.w);
#pragma unroll
    for (int c = 0; c < kStepCh; ++c)
      k.w.partials[((size_t)pslot * kNch + c) * a.nblocks + slot] = ch[c];
    const int f = sflags[0] | (sflags[1] << 16);
    if (!kDeferred) k.w.bflags[slot] = f;
    else if (f) atomicOr(k.w.mflags + 4 * mi + (int)(cur.step % 3), f);
    // this block alone reads its table entry: the next launch uses the
    // lists this search wrote
    if (mode == MODE_SEARCH) k.w.blocks[blockIdx.x].z = blk.z ^ kLpBit;
  }
  if (kDeferred && blk.w == 0) {
    // the member's first block carries the member forward: this step's
    // state for the next launch, and the ledger of the previous step (whose
    // partials and flags are complete now)
    if (threadIdx.x == 0) {
      BedDyn nd = cur;
      nd.fresh = 0;
      k.w.dyn[2 * mi + k.parity] = nd;
      k.w.mflags[4 * mi + (int)((cur.step + 1) % 3)] = 0;
    }
    if (!pd.fresh) bed_member_ledger<1, 1>(k, mi, pd.step, pflags, red);
  }
}

template <typename R, bool kDeferred>
__global__ __launch_bounds__(kThreads) void bed_step_kernel(BKArgs k)
{
  bed_step_body<R, kDeferred, false, false>(k);
}

template <typename R, bool kDeferred>
__global__ __launch_bounds__(kThreads) void bed_verlet_step_kernel(BKArgs k)
{
  bed_step_body<R, kDeferred, false, true>(k);
}

template <typename R>
__global__ __launch_bounds__(kThreads) void bed_velocity_kernel(BKArgs k)
{
  bed_step_body<R, false, true, true>(k);
}

// Energies of the current state on the current lists (no rebuild, no
// integration): kinetic, gravity, Hertz normal (2/5 f_el d) and tangential
// spring (s_t |xi|^2 / 2), in FP64.
__global__ __launch_bounds__(kThreads) void bed_energy_kernel(BKArgs k)
{
  __shared__ double red[4][kCh];
  __shared__ int sflags[2];
  const tdm_bed_arrays &a = k.a;
  const int4 blk = k.w.blocks[blockIdx.x];
  const int mi = blk.x, p0 = blk.y, cnt = blk.z & 0xFFFF, lp = (blk.z & kLpBit) ? 1 : 0;
  const bool valid = (int)threadIdx.x < cnt;
  const int p = p0 + (valid ? (int)threadIdx.x : 0);
  const int n = a.n;
  const size_t sn = (size_t)a.cand_slots * n;
  const BedConst m = load_const(a.members + mi);
  if (threadIdx.x < 2) sflags[threadIdx.x] = 0;
  double ch[4] = {0.0, 0.0, 0.0, 0.0};
  int status = 0;
  if (valid) {
    const double2 xp = k.x_in[p], vp = k.v_in[p], xb = k.w.xb[p];
    const double rp = a.radius[p];
    const double mass = m.rho_pi_t * (rp * rp);
    const uint32_t mask = k.w.mask[(size_t)lp * n + p];
    const int nc = k.w.ncand[(size_t)lp * n + p];
    const int *cand = k.w.cand + lp * sn + p;
    const double2 *xis = k.w.xi + lp * sn + p;
    const float2 *xis32 = reinterpret_cast<const float2 *>(k.w.xi) + lp * sn + p;
    const double ddx = xp.x - xb.x, ddy = xp.y - xb.y;
    const double lim = 0.5 * m.skin * 0.999;
    if (ddx * ddx + ddy * ddy >= lim * lim) status |= TDM_ST_STALE_LIST;
    ch[0] = 0.5 * mass * (vp.x * vp.x + vp.y * vp.y);
    ch[1] = -mass * m.gravity_y * xp.y;
    for (int e = 0; e < nc; ++e) {
      const int q = cand[(size_t)e * n];
      if (q < p) continue;
      const double2 xq = k.x_in[q];
      const double rq = a.radius[q];
      const double dx = xq.x - xp.x, dy = xq.y - xp.y;
      const double d = (rp + rq) - sqrt(dx * dx + dy * dy);
      if (!(d > 0)) continue;
      const double reff = (rp * rq) / (rp + rq);
      const double sn_ = (2 * m.e_star) * sqrt(reff * d);
      ch[2] += 0.4 * (((2.0 / 3.0) * sn_) * d) * d;
      if ((mask >> e) & 1u) {
        const double2 h = a.fp32_histories ? widen(xis32[(size_t)e * n]) : xis[(size_t)e * n];
        ch[3] += 0.5 * ((8 * m.g_star) * sqrt(reff * d)) * (h.x * h.x + h.y * h.y);
      }
    }
    const double gaps[3] = {xp.y - m.floor_y, xp.x - m.back_wall_x,
                            (m.right_x + m.right_v * m.time) - xp.x};
    const int use[3] = {m.use_floor, m.use_back_wall, m.use_right};
    for (int w = 0; w < 3; ++w) {
      const double d = rp - gaps[w];
      if (!use[w] || !(d > 0)) continue;
      const double sn_ = (2 * m.e_star) * sqrt(rp * d);
      ch[2] += 0.4 * (((2.0 / 3.0) * sn_) * d) * d;
      if ((mask >> (16 + w)) & 1u) {
        const double2 h = reinterpret_cast<const double2 *>(a.xi_wall)[3 * p + w];
        ch[3] += 0.5 * ((8 * m.g_star) * sqrt(rp * d)) * (h.x * h.x + h.y * h.y);
      }
    }
  }
  __syncthreads();
  if (status) atomicOr(&sflags[0], status);
  block_sum(ch, red);
  if (threadIdx.x == 0) {
    const size_t slot = (size_t)(m.block_first + blk.w);
#pragma unroll
    for (int c = 0; c < 4; ++c) k.w.partials[(size_t)(C_KIN + c) * a.nblocks + slot] = ch[c];
    k.w.bflags[slot] = sflags[0];
  }
}

// ---------------------------------------------------------------------------
// Ledger reductions.  A member's partials are reduced by one 256-thread
// block: thread t sums blocks t, t+256, t+512, ... in that order (loads
// issued four blocks at a time), then a fixed DPP tree per wave and the four
// wave totals in order.  The order depends only on the member's block count,
// so it is the same in the per-step finalize, in the next step's first block
// (multi-step calls) and in the flush at the end of a call, and independent
// of the batch around the member.
// ---------------------------------------------------------------------------
// U: blocks whose loads are issued together (the order of the sum does not
// depend on it; the step kernel uses 1 to keep its register budget).
// G: 256-thread groups of the block (blockDim = 256 G); group g takes the
// channels [g NC/G, (g+1) NC/G) with the same thread-to-block assignment as a
// single group, so the sums are the same for every G.
template <int NC, int U = 4, int G = 1>
__device__ void bed_reduce(const BKArgs &k, int b0, int nb, int pslot, int c_lo,
                           double (*wsum)[kCh], double *stot)
{
  static_assert(NC % G == 0, "channels split evenly over the groups");
  constexpr int NG = NC / G;
  const int g = G == 1 ? 0 : (int)threadIdx.x / kThreads;
  const int t = G == 1 ? (int)threadIdx.x : (int)threadIdx.x % kThreads;
  double acc[NG];
#pragma unroll
  for (int c = 0; c < NG; ++c) acc[c] = 0.0;
  const double *part =
      k.w.partials + ((size_t)pslot * kNch + c_lo + g * NG) * k.a.nblocks + b0;
  for (int base = 0; base < nb; base += U * kThreads) {
    double v[U][NG];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int b = base + u * kThreads + t;
#pragma unroll
      for (int c = 0; c < NG; ++c) v[u][c] = b < nb ? part[(size_t)c * k.a.nblocks + b] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < U; ++u)
#pragma unroll
      for (int c = 0; c < NG; ++c) acc[c] += v[u][c];
  }
  const int lane = t & 63, wave = t >> 6;
#pragma unroll
  for (int c = 0; c < NG; ++c) {
    double tot = 0.0;
    if (__ballot(acc[c] != 0.0)) tot = wave_sum(acc[c]);
    if (lane == 0) wsum[wave][g * NG + c] = tot;
  }
  __syncthreads();
  if (threadIdx.x < NC)
    stot[threadIdx.x] = ((wsum[0][threadIdx.x] + wsum[1][threadIdx.x]) + wsum[2][threadIdx.x]) +
                        wsum[3][threadIdx.x];
  __syncthreads();
}

// Clock, search phase and list parity after a step that started in `d` and
// raised `flags` (status | rebuild request << 16).
__device__ inline BedDyn bed_advance(const BedDyn &d, int flags, double dt)
{
  BedDyn r = d;
  r.time = d.time + dt;
  r.step = d.step + 1;
  r.fresh = 0;
  if (d.mode == MODE_SEARCH) {
    r.lp = d.lp ^ 1;
    r.rebuilds = d.rebuilds + 1;
  }
  if (d.mode == MODE_BIN) { r.tab = d.tab ^ 1; r.mode = MODE_SEARCH; }
  else r.mode = (flags >> 16) ? MODE_BIN : MODE_IDLE;
  return r;
}

// Step ledger of member mi: the partials of the step that started at clock
// d.step (parity slot d.step & 1) are reduced and added to the accounting;
// status bits are merged into the member record.  Called by every thread of
// a block of 256 G threads; thread 0 writes.
template <int U = 4, int G = 1>
__device__ void bed_member_ledger(const BKArgs &k, int mi, long long step, int flags,
                                  double (*wsum)[kCh])
{
  __shared__ double stot[kStepCh];
  tdm_bed_member *mp = k.a.members + mi;
  const int b0 = mp->block_first, nb = mp->nblocks;
  __syncthreads();  // wsum may be the caller's reduction scratch
  bed_reduce<kStepCh, U, G>(k, b0, nb, (int)(step & 1), 0, wsum, stot);
  if (threadIdx.x != 0) return;
  double *acc = k.a.accounting + (size_t)mi * TDM_BED_NACC;
  const double dt = mp->dt;
  acc[TDM_BED_DASHPOT] = acc[TDM_BED_DASHPOT] + dt * stot[C_DASH];
  acc[TDM_BED_FRICTION] = acc[TDM_BED_FRICTION] + dt * stot[C_FRIC];
  acc[TDM_BED_WALL_DASHPOT] = acc[TDM_BED_WALL_DASHPOT] + dt * stot[C_WDASH];
  acc[TDM_BED_WALL_FRICTION] = acc[TDM_BED_WALL_FRICTION] + dt * stot[C_WFRIC];
  acc[TDM_BED_WALL_IMPULSE] = acc[TDM_BED_WALL_IMPULSE] + dt * stot[C_WFX];
  acc[TDM_BED_WALL_IMPULSE + 1] = acc[TDM_BED_WALL_IMPULSE + 1] + dt * stot[C_WFY];
  const double nwall = floor(stot[C_COUNTS] / kWallCount);
  acc[TDM_BED_CONTACTS] = stot[C_COUNTS] - kWallCount * nwall;
  acc[TDM_BED_WALL_CONTACTS] = nwall;
  acc[TDM_BED_WALL_WORK] = acc[TDM_BED_WALL_WORK] + dt * stot[C_WWORK];
  if (flags & 0xFFFF) atomicOr(&mp->status, flags & 0xFFFF);
}

__device__ inline void bed_write_record(tdm_bed_member *mp, const BedDyn &d)
{
  mp->time = d.time;
  mp->step_count = d.step;
  mp->mode = d.mode;
  mp->tab = d.tab;
  mp->list_parity = d.lp;
  mp->rebuilds = d.rebuilds;
  mp->verlet_start = 0;  // a step has been taken from the state
}

__device__ inline BedDyn bed_record_state(const tdm_bed_member *mp)
{
  BedDyn d;
  d.time = mp->time;
  d.step = mp->step_count;
  d.mode = mp->mode;
  d.tab = mp->tab;
  d.lp = mp->list_parity;
  d.rebuilds = mp->rebuilds;
  d.fresh = 1;
  d.pad = 0;
  return d;
}

// Finalize of a single launch (one step or an energy evaluation): one block
// per member.
__global__ __launch_bounds__(kThreads) void bed_finalize_kernel(BKArgs k)
{
  __shared__ int sflag;
  __shared__ double wsum[4][kCh];
  __shared__ double stot[4];
  const tdm_bed_arrays &a = k.a;
  const int mi = blockIdx.x;
  tdm_bed_member *mp = a.members + mi;
  const int b0 = mp->block_first, nb = mp->nblocks;
  if (threadIdx.x == 0) sflag = 0;
  __syncthreads();
  int fl = 0;
  for (int b = threadIdx.x; b < nb; b += kThreads) fl |= k.w.bflags[b0 + b];
  if (fl) atomicOr(&sflag, fl);
  __syncthreads();
  const int flags = sflag;
  if (k.op == OP_ENERGY) {
    bed_reduce<4>(k, b0, nb, 0, C_KIN, wsum, stot);
    if (threadIdx.x != 0) return;
    double *en = a.energy + (size_t)mi * 4;
#pragma unroll
    for (int c = 0; c < 4; ++c) en[c] = stot[c];
    if (flags & 0xFFFF) mp->status |= flags & 0xFFFF;
    return;
  }
  const BedDyn d = bed_record_state(mp);
  // (a search step's blocks flipped their own table entries' list parity)
  bed_member_ledger(k, mi, 0, flags, wsum);
  if (threadIdx.x == 0) {
    bed_write_record(mp, bed_advance(d, flags, mp->dt));
    if (mp->integrator && !k.a.verlet) mp->status |= TDM_ST_INTEGRATOR;
  }
}

// End of a multi-step call: ledger of the last step, the member record
// brought to the state after it, and the member's flag words cleared for the
// next call (its first launch raises flags into a word it does not clear).
// k.parity is the last launch's parity.  Four 256-thread groups split the
// channels, so the partials of up to 16 x 256 blocks are loaded in one round.
constexpr int kFlushGroups = 4;
constexpr int kFlushU = 16;
__global__ __launch_bounds__(kThreads * kFlushGroups) void bed_flush_kernel(BKArgs k)
{
  __shared__ double wsum[4][kCh];
  const int mi = blockIdx.x;
  const BedDyn st = k.w.dyn[2 * mi + k.parity];
  const int f = k.w.mflags[4 * mi + (int)(st.step % 3)];
  bed_member_ledger<kFlushU, kFlushGroups>(k, mi, st.step, f, wsum);
  if (threadIdx.x == 0) {
    tdm_bed_member *mp = k.a.members + mi;
    bed_write_record(mp, bed_advance(st, f, mp->dt));
    if (mp->integrator && !k.a.verlet) mp->status |= TDM_ST_INTEGRATOR;
    reinterpret_cast<int4 *>(k.w.mflags)[mi] = make_int4(0, 0, 0, 0);
  }
}

__global__ void bed_bin_kernel(tdm_bed_arrays a, BedWs w)
{
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.n) return;
  int lo = 0, hi = a.nmembers - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (a.members[mid].first <= t) lo = mid; else hi = mid - 1;
  }
  const tdm_bed_member &m = a.members[lo];
  int cx, cy;
  const int cell = cell_of(m.grid_origin[0], m.grid_origin[1], m.cell_size, m.ncx, m.ncy,
                           m.cell_first, reinterpret_cast<const double2 *>(a.x)[t], &cx, &cy);
  const int slot = atomicAdd(w.cell_count + (size_t)m.tab * a.ncells + cell, 1);
  if (slot < a.cell_cap) w.cell_tab[((size_t)m.tab * a.ncells + cell) * a.cell_cap + slot] = t;
  else atomicOr(&a.members[lo].status, TDM_ST_CELL_OVERFLOW);
}

__global__ void bed_set_mode_kernel(tdm_bed_member *m, int nm, int mode)
{
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < nm) m[t].mode = mode;
}

int check_bed(const tdm_bed_arrays *a, BedWs *w)
{
  if (!a || !a->members || !a->x || !a->v || !a->x_alt || !a->v_alt || !a->radius ||
      !a->xi_wall || !a->accounting || !a->workspace)
    return fail(TDM_EINVAL, "null array in tdm_bed_arrays");
  if (a->n <= 0 || a->nmembers <= 0 || a->nblocks <= 0 || a->ncells <= 0)
    return fail(TDM_EINVAL, "empty bed batch");
  if (a->cand_slots <= 0 || a->cand_slots > kMaxSlots || a->cell_cap <= 0)
    return fail(TDM_EINVAL, "cand_slots must be in [1,16] and cell_cap > 0");
  if ((long long)a->ncells * a->cell_cap >= (1LL << 31))
    return fail(TDM_EINVAL, "bed cell table too large (ncells x cell_cap must stay below 2^31)");
  size_t need = 0;
  carve(a->workspace, a->n, a->nblocks, a->nmembers, a->ncells, a->cand_slots, a->cell_cap, w,
        &need);
  if (need > a->workspace_bytes) return fail(TDM_ESPACE, "bed workspace too small");
  if (reinterpret_cast<uintptr_t>(a->workspace) & 255)
    return fail(TDM_EINVAL, "workspace must be 256-byte aligned");
  return TDM_OK;
}

BKArgs bed_args(const tdm_bed_arrays *a, const BedWs &w, int parity, int op, int first = 0)
{
  BKArgs k;
  k.a = *a;
  k.w = w;
  k.op = op;
  k.parity = parity;
  k.first = first;
  const double2 *x = reinterpret_cast<const double2 *>(a->x);
  const double2 *v = reinterpret_cast<const double2 *>(a->v);
  const double2 *xa = reinterpret_cast<const double2 *>(a->x_alt);
  const double2 *va = reinterpret_cast<const double2 *>(a->v_alt);
  k.x_in = parity ? xa : x;
  k.v_in = parity ? va : v;
  k.x_out = const_cast<double2 *>(parity ? x : xa);
  k.v_out = const_cast<double2 *>(parity ? v : va);
  return k;
}

// One launch of the step (or energy) kernel; a single launch (deferred = 0)
// is followed by its finalize.  first: the first launch of a multi-step call.
int launch_bed(const tdm_bed_arrays *a, const BedWs &w, int parity, int fp32, int op,
               hipStream_t s, hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr,
               int deferred = 0, int first = 0)
{
  const BKArgs k = bed_args(a, w, parity, op, first);
  if (op == OP_ENERGY) {
    bed_energy_kernel<<<a->nblocks, kThreads, 0, s>>>(k);
  } else {
    void (*kern)(BKArgs) =
        a->verlet
            ? (deferred ? (fp32 ? bed_verlet_step_kernel<float, true>
                                : bed_verlet_step_kernel<double, true>)
                        : (fp32 ? bed_verlet_step_kernel<float, false>
                                : bed_verlet_step_kernel<double, false>))
            : (deferred ? (fp32 ? bed_step_kernel<float, true> : bed_step_kernel<double, true>)
                        : (fp32 ? bed_step_kernel<float, false> : bed_step_kernel<double, false>));
    if (ev0) {
      // events carried by the dispatch itself: they time exactly this kernel
      hipExtLaunchKernelGGL(kern, dim3(a->nblocks), dim3(kThreads), 0, s, ev0, ev1, 0, k);
    } else {
      kern<<<a->nblocks, kThreads, 0, s>>>(k);
    }
  }
  HIP_TRY(hipGetLastError());
  if (!deferred) {
    bed_finalize_kernel<<<a->nmembers, kThreads, 0, s>>>(k);
    HIP_TRY(hipGetLastError());
  }
  return TDM_OK;
}

// Batches of at most this many members carry the ledger from launch to
// launch inside a multi-step call; larger batches keep one finalize launch
// per step (one reduction block per member in its own launch is cheaper
// there than a reduction inside the first block of every member).
constexpr int kDeferMaxMembers = 4;

// nsteps steps from parity 0: one launch per step.  For two or more steps of
// a small batch the per-step finalize launches are replaced by the deferred
// ledger, closed by a flush launch.  ev (optional) holds 2 events per step
// around each step launch.
int launch_bed_steps(const tdm_bed_arrays *a, const BedWs &w, int nsteps, int fp32,
                     hipStream_t s, hipEvent_t *ev = nullptr)
{
  if (nsteps <= 0) return TDM_OK;
  int parity = 0;
  if (nsteps == 1 || a->nmembers > kDeferMaxMembers) {
    for (int i = 0; i < nsteps; ++i) {
      int rc = launch_bed(a, w, parity, fp32, OP_STEP, s, ev ? ev[2 * i] : nullptr,
                          ev ? ev[2 * i + 1] : nullptr);
      if (rc) return rc;
      parity ^= 1;
    }
    return TDM_OK;
  }
  for (int i = 0; i < nsteps; ++i) {
    int rc = launch_bed(a, w, parity, fp32, OP_STEP, s, ev ? ev[2 * i] : nullptr,
                        ev ? ev[2 * i + 1] : nullptr, 1, i == 0);
    if (rc) return rc;
    parity ^= 1;
  }
  const BKArgs k = bed_args(a, w, parity ^ 1, OP_STEP);
  bed_flush_kernel<<<a->nmembers, kThreads * kFlushGroups, 0, s>>>(k);
  HIP_TRY(hipGetLastError());
  return TDM_OK;
}

}  // namespace

extern "C" {

int tdm_bed_abi_sizes(size_t *member_bytes, size_t *arrays_bytes)
{
  if (!member_bytes || !arrays_bytes) return fail(TDM_EINVAL, "null output");
  *member_bytes = sizeof(tdm_bed_member);
  *arrays_bytes = sizeof(tdm_bed_arrays);
  return TDM_OK;
}

int tdm_bed_blocks(int n)
{
  if (n <= 0) return fail(TDM_EINVAL, "bed member needs n > 0");
  return (n + kThreads - 1) / kThreads;
}

size_t tdm_bed_workspace_bytes(int n, int nblocks, int nmembers, int ncells, int cand_slots,
                               int cell_cap)
{
  BedWs w;
  size_t total = 0;
  carve(nullptr, n, nblocks, nmembers, ncells, cand_slots, cell_cap, &w, &total);
  return total;
}

int tdm_bed_workspace_layout(int n, int nblocks, int nmembers, int ncells, int cand_slots,
                             int cell_cap, size_t *cand_off, size_t *xi_off, size_t *mask_off,
                             size_t *count_off, size_t *xb_off)
{
  if (!cand_off || !xi_off || !mask_off || !count_off || !xb_off)
    return fail(TDM_EINVAL, "null output");
  BedWs w;
  size_t total = 0;
  carve(nullptr, n, nblocks, nmembers, ncells, cand_slots, cell_cap, &w, &total);
  *cand_off = reinterpret_cast<size_t>(w.cand);
  *xi_off = reinterpret_cast<size_t>(w.xi);
  *mask_off = reinterpret_cast<size_t>(w.mask);
  *count_off = reinterpret_cast<size_t>(w.ncand);
  *xb_off = reinterpret_cast<size_t>(w.xb);
  return TDM_OK;
}

int tdm_bed_setup(const tdm_bed_arrays *a, void *stream)
{
  BedWs w;
  int rc = check_bed(a, &w);
  if (rc) return rc;
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int nm = a->nmembers;
  tdm_bed_member *hm = new tdm_bed_member[nm];
  hipError_t e = hipMemcpyAsync(hm, a->members, sizeof(tdm_bed_member) * nm,
                                hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) { delete[] hm; return fail(TDM_EHIP, hipGetErrorString(e)); }
  int total = 0;
  for (int i = 0; i < nm; ++i) {
    if (hm[i].block_first != total || hm[i].nblocks != tdm_bed_blocks(hm[i].n)) {
      delete[] hm;
      return fail(TDM_EINVAL, "bed member block_first/nblocks do not match tdm_bed_blocks");
    }
    if ((hm[i].integrator != 0 && hm[i].integrator != 1) || (hm[i].integrator && !a->verlet)) {
      delete[] hm;
      return fail(TDM_EINVAL, "bed member integrator must be 0 or 1, and a batch with "
                              "velocity-Verlet members needs tdm_bed_arrays.verlet");
    }
    total += hm[i].nblocks;
  }
  if (total != a->nblocks) { delete[] hm; return fail(TDM_EINVAL, "bed nblocks mismatch"); }
  // Block table: launch slot -> (member, first disk, count | list parity,
  // block index within the member).  Partials are indexed by the member's
  // own block index, so the launch order is free: blocks are dealt
  // round-robin over the 8 XCDs (slot s runs on XCD s % 8), and a disk
  // gathers its partners from the member's other blocks, so a member of at
  // most kXcdWhole blocks runs entirely on one XCD (the one with the fewest
  // blocks so far) and a larger one in 8 contiguous runs, one per XCD.
  constexpr int kXcd = 8, kXcdWhole = 32;
  std::vector<int2> queue[kXcd];  // (member, block index)
  for (int i = 0; i < nm; ++i) {
    const int nb = hm[i].nblocks;
    if (nb <= kXcdWhole) {
      int x = 0;
      for (int c = 1; c < kXcd; ++c)
        if (queue[c].size() < queue[x].size()) x = c;
      for (int b = 0; b < nb; ++b) queue[x].push_back(make_int2(i, b));
    } else {
      for (int r = 0; r < kXcd; ++r)
        for (int b = r * nb / kXcd; b < (r + 1) * nb / kXcd; ++b)
          queue[r].push_back(make_int2(i, b));
    }
  }
  int4 *hb = new int4[total];
  size_t head[kXcd] = {};
  for (int slot = 0; slot < total; ++slot) {
    int x = slot % kXcd;
    if (head[x] == queue[x].size()) {
      // this XCD's share is spent: take the next block of the XCD with the most left
      for (int c = 0; c < kXcd; ++c)
        if (queue[c].size() - head[c] > queue[x].size() - head[x]) x = c;
    }
    const int2 it = queue[x][head[x]++];
    const int i = it.x, item = it.y;
    const int first = hm[i].first + item * kThreads;
    const int count = min(kThreads, hm[i].first + hm[i].n - first);
    hb[slot] = make_int4(i, first, count | (hm[i].list_parity ? kLpBit : 0), item);
  }
  e = hipMemcpyAsync(w.blocks, hb, sizeof(int4) * total, hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  delete[] hb;
  delete[] hm;
  if (e != hipSuccess) return fail(TDM_EHIP, hipGetErrorString(e));
  fill_kernel<int><<<64, 256, 0, s>>>(w.bflags, (size_t)a->nblocks, 0);
  fill_kernel<int><<<64, 256, 0, s>>>(w.mflags, 4 * (size_t)a->nmembers, 0);
  fill_kernel<uint32_t><<<1024, 256, 0, s>>>(w.mask, 2 * (size_t)a->n, 0u);
  fill_kernel<double><<<1024, 256, 0, s>>>(a->xi_wall, 6 * (size_t)a->n, 0.0);
  HIP_TRY(hipGetLastError());
  return tdm_bed_rebin(a, stream);
}

int tdm_bed_rebin(const tdm_bed_arrays *a, void *stream)
{
  BedWs w;
  int rc = check_bed(a, &w);
  if (rc) return rc;
  hipStream_t s = static_cast<hipStream_t>(stream);
  fill_kernel<int><<<1024, 256, 0, s>>>(w.cell_count, 2 * (size_t)a->ncells, 0);
  bed_bin_kernel<<<(a->n + 255) / 256, 256, 0, s>>>(*a, w);
  bed_set_mode_kernel<<<(a->nmembers + 255) / 256, 256, 0, s>>>(a->members, a->nmembers,
                                                                 MODE_SEARCH);
  HIP_TRY(hipGetLastError());
  return TDM_OK;
}

int tdm_bed_step(const tdm_bed_arrays *a, int nsteps, int fp32, void *stream)
{
  BedWs w;
  int rc = check_bed(a, &w);
  if (rc) return rc;
  if ((fp32 != 0) != (a->fp32_histories != 0))
    return fail(TDM_EINVAL, "the bed's spring histories are stored in the other precision "
                            "(fp32_histories must match the step's fp32 flag)");
  if (nsteps < 0) return fail(TDM_EINVAL, "bad nsteps");
  hipStream_t s = static_cast<hipStream_t>(stream);
  return launch_bed_steps(a, w, nsteps, fp32, s);
}

int tdm_bed_step_timed(const tdm_bed_arrays *a, int nsteps, int fp32, void *stream,
                       double *mean_kernel_us)
{
  BedWs w;
  int rc = check_bed(a, &w);
  if (rc) return rc;
  if ((fp32 != 0) != (a->fp32_histories != 0))
    return fail(TDM_EINVAL, "the bed's spring histories are stored in the other precision "
                            "(fp32_histories must match the step's fp32 flag)");
  if (nsteps <= 0 || (nsteps & 1) || !mean_kernel_us) return fail(TDM_EINVAL, "bad timing request");
  hipStream_t s = static_cast<hipStream_t>(stream);
  hipEvent_t *ev = new hipEvent_t[2 * nsteps]();
  int made = 0;
  for (; made < 2 * nsteps; ++made)
    if (hipEventCreate(&ev[made]) != hipSuccess) break;
  if (made < 2 * nsteps) rc = fail(TDM_EHIP, "hipEventCreate failed");
  if (rc == TDM_OK) rc = launch_bed_steps(a, w, nsteps, fp32, s, ev);
  double total = 0.0;
  if (rc == TDM_OK) {
    hipError_t e = hipStreamSynchronize(s);
    for (int i = 0; i < nsteps && e == hipSuccess; ++i) {
      float ms = 0.0f;
      e = hipEventElapsedTime(&ms, ev[2 * i], ev[2 * i + 1]);
      total += ms;
    }
    if (e != hipSuccess) rc = fail(TDM_EHIP, hipGetErrorString(e));
  }
  for (int i = 0; i < made; ++i) (void)hipEventDestroy(ev[i]);
  delete[] ev;
  if (rc == TDM_OK) *mean_kernel_us = total * 1e3 / nsteps;
  return rc;
}

int tdm_bed_energy(const tdm_bed_arrays *a, void *stream)
{
  BedWs w;
  int rc = check_bed(a, &w);
  if (rc) return rc;
  if (!a->energy) return fail(TDM_EINVAL, "energy output required");
  return launch_bed(a, w, 0, 0, OP_ENERGY, static_cast<hipStream_t>(stream));
}

int tdm_bed_velocity(const tdm_bed_arrays *a, int fp32, void *stream)
{
  BedWs w;
  int rc = check_bed(a, &w);
  if (rc) return rc;
  if ((fp32 != 0) != (a->fp32_histories != 0))
    return fail(TDM_EINVAL, "the bed's spring histories are stored in the other precision "
                            "(fp32_histories must match the fp32 flag)");
  // parity 0: reads x, v, writes v_alt
  const BKArgs k = bed_args(a, w, 0, OP_STEP);
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (fp32) bed_velocity_kernel<float><<<a->nblocks, kThreads, 0, s>>>(k);
  else bed_velocity_kernel<double><<<a->nblocks, kThreads, 0, s>>>(k);
  HIP_TRY(hipGetLastError());
  return TDM_OK;
}

int tdm_bed_graph_create(const tdm_bed_arrays *a, int pairs, int fp32, void *stream,
                         tdm_graph **out)
{
  BedWs w;
  int rc = check_bed(a, &w);
  if (rc) return rc;
  if ((fp32 != 0) != (a->fp32_histories != 0))
    return fail(TDM_EINVAL, "the bed's spring histories are stored in the other precision "
                            "(fp32_histories must match the step's fp32 flag)");
  if (!out || pairs <= 0) return fail(TDM_EINVAL, "bad graph request");
  hipStream_t s = static_cast<hipStream_t>(stream);
  HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  rc = launch_bed_steps(a, w, 2 * pairs, fp32, s);
  hipGraph_t g;
  hipError_t e = hipStreamEndCapture(s, &g);
  if (rc) return rc;
  if (e != hipSuccess) return fail(TDM_EHIP, hipGetErrorString(e));
  hipGraphExec_t ex;
  e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
  if (e != hipSuccess) { (void)hipGraphDestroy(g); return fail(TDM_EHIP, hipGetErrorString(e)); }
  *out = new tdm_graph{g, ex};
  return TDM_OK;
}

}  // extern "C"
