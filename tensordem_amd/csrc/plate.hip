// Fused force-and-integrate step for bonded ice plates on MI355X (gfx950).
//
// One launch advances every member (plate) of a batch by one step of the
// reference solver IceDEM.step (src/tensordem/dem.py:347-382):
//   * each 256-thread block owns a TX x TY tile of one lattice and stages the
//     tile plus a 2-disk halo of positions in LDS;
//   * the objective shear strain (dem.py:182-208) of the tile and a 1-disk
//     ring is computed from LDS, so it never round-trips through HBM;
//   * every disk gathers its bonds (8-neighbour stencil, no index lists) and
//     its unbonded contact candidates in the reference's summation order, so
//     forces are deterministic without float atomics;
//   * fracture is committed by both ends of a bond with identical arithmetic
//     (the lower-index end writes the bond's failure record);
//   * tool or hull contact, loads, semi-implicit Euler, fixed clamp and the
//     per-block ledger partials are fused in; a per-member finalize launch
//     sums that member's partials in a fixed order and advances its clock;
//   * the cell hash for unbonded contacts is rebuilt in two pipelined phases
//     (bin, then search) folded into ordinary steps, triggered on device when
//     a disk has moved a quarter of the Verlet skin.
#include "common.h"

#include <vector>

namespace {

using namespace tdm;

constexpr int kHalo = 2;
constexpr int kMaxLds = 1300;   // (TX+4)*(TY+4) for TX in 16..256, TY = 256/TX
constexpr int kMaxRing = 800;   // (TX+2)*(TY+2)
constexpr int kMaxCand = 16;    // upper bound for cand_slots
constexpr int kScan = 4;        // entries gathered per round (candidate search, candidate forces)
constexpr int kCandLds = 6;     // candidate slots whose pair forces are staged in LDS
// staged kernels serve tiles at most 128 wide, which bounds their halo and
// ring (132 x 6, 130 x 4) and keeps an FP32 block within 40 KB of LDS: four
// blocks per compute unit, as without staging
constexpr int kStageMaxTile = 128;
constexpr int kMaxLdsStage = (kStageMaxTile + 4) * (kThreads / kStageMaxTile + 4);
constexpr int kMaxRingStage = (kStageMaxTile + 2) * (kThreads / kStageMaxTile + 2);
constexpr int kCandBatch = 2;   // candidates gathered per round of the force pre-pass

// ledger partial channels
enum {
  C_RELEASED = 0, C_ACQUIRED, C_PAIR_DISS, C_TOOL_DISS, C_VSQ, C_TOOL_FX, C_TOOL_FY,
  C_SUP_X, C_SUP_Y, C_VSUM_X, C_VSUM_Y, C_LOAD_X, C_LOAD_Y, C_WORK, C_MOMENT, C_BREAKS
};

enum { MODE_IDLE = 0, MODE_SEARCH = 1, MODE_BIN = 2 };
enum { OP_FORCES = 0, OP_STEP = 1, OP_ENERGY = 2 };
// energy channels (OP_ENERGY reuses the partial buffer)
enum { E_BOND = 0, E_CONTACT, E_TOOL, E_VSQ };

// direction codes in ascending flat-index offset: -nx-1, -nx, -nx+1, -1, +1, nx-1, nx, nx+1
__host__ __device__ constexpr int dxc(int c) { return c < 3 ? c - 1 : (c == 3 ? -1 : (c == 4 ? 1 : c - 6)); }
__host__ __device__ constexpr int dyc(int c) { return c < 3 ? -1 : (c < 5 ? 0 : 1); }

// Clock and contact-search phase of a member at the start of a step, carried
// from launch to launch inside one multi-step call (double-buffered by the
// launch parity); the member record is brought up to date when the call ends.
struct MemberDyn {
  double time;
  long long step;
  int mode, tab, fresh, pad;
};

struct Workspace {
  int4 *blocks;          // 2 per block: (member, tile x0, tile y0, rank), (nx, ny, first, tile_x)
  double *partials;      // [2][kCh][nblocks], channel-major, by step parity
  int *bflags;           // [nblocks] status bits | rebuild request << 16
  int *cell_count;       // [2][ncells]
  int *cell_tab;         // [2][ncells][cap]
  int *cand;             // [slots][n]
  uint8_t *ncand;        // [n]
  float2 *xb;            // [n] position - lattice site when the list was built
  MemberDyn *dyn;        // [nmembers][2]
  int *mflags;           // [nmembers][4]: status | request << 16 of steps s % 3
};


__host__ __device__ inline void carve(void *base, int n, int nblocks, int nmembers, int ncells,
                                      int slots, int cap, Workspace *w, size_t *total)
{
  size_t off = 0;
  char *b = static_cast<char *>(base);
  auto take = [&](size_t bytes) {
    char *p = b ? b + off : nullptr;
    off += align256(bytes);
    return p;
  };
  w->blocks = reinterpret_cast<int4 *>(take(2 * sizeof(int4) * size_t(nblocks)));
  w->partials = reinterpret_cast<double *>(take(sizeof(double) * 2 * kCh * size_t(nblocks)));
  w->bflags = reinterpret_cast<int *>(take(sizeof(int) * size_t(nblocks)));
  w->cell_count = reinterpret_cast<int *>(take(sizeof(int) * 2 * size_t(ncells)));
  w->cell_tab = reinterpret_cast<int *>(take(sizeof(int) * 2 * size_t(ncells) * size_t(cap)));
  w->cand = reinterpret_cast<int *>(take(sizeof(int) * size_t(slots) * size_t(n)));
  w->ncand = reinterpret_cast<uint8_t *>(take(size_t(n)));
  w->xb = reinterpret_cast<float2 *>(take(sizeof(float2) * size_t(n)));
  w->dyn = reinterpret_cast<MemberDyn *>(take(sizeof(MemberDyn) * 2 * size_t(nmembers)));
  w->mflags = reinterpret_cast<int *>(take(sizeof(int) * 4 * size_t(nmembers)));
  *total = off;
}


__device__ inline double2 lattice_site(int ix, int iy, double half_span, double two_r)
{
  // dem.py:105-110: (ix - (nx-1)/2) * 2r, iy * 2r
  return make_double2(((double)ix - half_span) * two_r, (double)iy * two_r);
}

__device__ inline bool on_lattice(int ix, int iy, int nx, int ny)
{
  return ix >= 0 && ix < nx && iy >= 0 && iy < ny;
}

__device__ inline uint32_t present_links(int ix, int iy, int nx, int ny)
{
  uint32_t m = 0;
#pragma unroll
  for (int c = 0; c < 8; ++c)
    if (on_lattice(ix + dxc(c), iy + dyc(c), nx, ny)) m |= 1u << c;
  return m;
}

// ---------------------------------------------------------------------------
// Shear strain (dem.py:182-208).  Bonds are visited as index_add_ visits
// them: partners above the disk by ascending index (codes 4..7), then
// partners below (codes 0..3).  pinv of the 2x2 covariance is the closed-form
// inverse.
// ---------------------------------------------------------------------------
struct Tile {
  const double2 *sx;  // LDS positions, row stride lw, origin (tx0-2, ty0-2)
  int lw, tx0, ty0, nx, ny;
  double half_span, two_r;
  __device__ double2 x(int ix, int iy) const { return sx[(iy - ty0 + kHalo) * lw + (ix - tx0 + kHalo)]; }
};

__device__ double strain64(const Tile &t, int ix, int iy)
{
  double c00 = 0, c01 = 0, c10 = 0, c11 = 0, x00 = 0, x01 = 0, x10 = 0, x11 = 0;
  const double2 xp = t.x(ix, iy);
  const double2 sp = lattice_site(ix, iy, t.half_span, t.two_r);
#pragma unroll
  for (int k = 0; k < 8; ++k) {
    const int c = (k + 4) & 7;
    const int jx = ix + dxc(c), jy = iy + dyc(c);
    if (!on_lattice(jx, jy, t.nx, t.ny)) continue;
    const double2 xq = t.x(jx, jy);
    const double2 sq = lattice_site(jx, jy, t.half_span, t.two_r);
    double rx, ry, qx, qy;
    if (c >= 4) { rx = sq.x - sp.x; ry = sq.y - sp.y; qx = xq.x - xp.x; qy = xq.y - xp.y; }
    else        { rx = sp.x - sq.x; ry = sp.y - sq.y; qx = xp.x - xq.x; qy = xp.y - xq.y; }
    c00 += rx * rx; c01 += rx * ry; c10 += ry * rx; c11 += ry * ry;
    x00 += qx * rx; x01 += qx * ry; x10 += qy * rx; x11 += qy * ry;
  }
  double det = c00 * c11 - c01 * c10;
  double i00, i01, i10, i11;
  if (det != 0.0) {
    i00 = c11 / det; i01 = -c01 / det; i10 = -c10 / det; i11 = c00 / det;
  } else {
    double tr = c00 + c11;
    if (tr > 0.0) {
      double t2 = tr * tr;
      i00 = c00 / t2; i01 = c01 / t2; i10 = c10 / t2; i11 = c11 / t2;
    } else {
      i00 = i01 = i10 = i11 = 0.0;
    }
  }
  double f00 = x00 * i00 + x01 * i10;
  double f01 = x00 * i01 + x01 * i11;
  double f10 = x10 * i00 + x11 * i10;
  double f11 = x10 * i01 + x11 * i11;
  double g00 = f00 * f00 + f10 * f10;
  double g01 = f00 * f01 + f10 * f11;
  double g11 = f01 * f01 + f11 * f11;
  double exx = 0.5 * (g00 - 1.0), exy = 0.5 * (g01 - 0.0), eyy = 0.5 * (g11 - 1.0);
  double dxy = exx - eyy;
  return sqrt(dxy * dxy + 4.0 * (exy * exy));
}

struct KArgs {
  tdm_plate_arrays a;
  Workspace w;
  const double2 *x_in, *v_in;
  double2 *x_out, *v_out;
  int op;        // OP_FORCES, OP_STEP or OP_ENERGY
  int parity;    // 0: x, v -> x_alt, v_alt; 1: the reverse
  int deferred;  // OP_STEP inside a multi-step call: clock from `dyn`, ledger one launch later
  int first;     // the first launch of a multi-step call: the member records hold the clock
};


// Repulsive contact of a non-bonded or broken pair (dem.py:278-288), written
// with the lo -> hi orientation so both ends compute identical bits.  Returns
// false when the pair does not overlap; else (Fx, Fy) is the force on lo's
// partner side convention of the reference (`pair_force`, applied +F to lo,
// -F to hi) and `diss` the pair's dissipated power (dem.py:327-328).  The
// normal of coincident disks falls back to the pair's unit rest vector
// (dem.py:241-242), which n0() forms only in that case.
template <typename R, typename F, typename N0>
__device__ inline bool contact_pair(double2 x_lo, double2 x_hi, double2 v_lo, double2 v_hi,
                                    N0 &&n0, double radius, double kc, double cdamp,
                                    F &Fx, F &Fy, double &diss)
{
  const double two_r = 2 * radius;
  if constexpr (Prec<R>::fp64) {
    double dx = x_hi.x - x_lo.x, dy = x_hi.y - x_lo.y;
    double d = sqrt(dx * dx + dy * dy);
    if (!(d < two_r)) return false;
    double nx, ny;
    if (d > 0) { double dd = d > kEps64 ? d : kEps64; nx = dx / dd; ny = dy / dd; }
    else { const double2 u = n0(); nx = u.x; ny = u.y; }
    double ov = two_r - d;
    double vn = (v_hi.x - v_lo.x) * nx + (v_hi.y - v_lo.y) * ny;
    double rep = kc * ov - cdamp * vn;
    if (rep < 0) rep = 0;
    double mag = -rep;
    Fx = mag * nx; Fy = mag * ny;
    diss = (rep - kc * ov) * (-vn);
  } else {
    float dx = (float)(x_hi.x - x_lo.x), dy = (float)(x_hi.y - x_lo.y);
    float d = sqrt32(dx * dx + dy * dy);
    float tr = (float)two_r;
    if (!(d < tr)) return false;
    float nx, ny;
    if (d > 0) { float r = rcp32(fmaxf(d, kEps32)); nx = dx * r; ny = dy * r; }
    else { const double2 u = n0(); nx = (float)u.x; ny = (float)u.y; }
    float ov = tr - d;
    float vn = (float)(v_hi.x - v_lo.x) * nx + (float)(v_hi.y - v_lo.y) * ny;
    float kcf = (float)kc;
    float rep = fmaxf(kcf * ov - (float)cdamp * vn, 0.0f);
    Fx = (F)(-rep * nx); Fy = (F)(-rep * ny);
    diss = (double)((rep - kcf * ov) * (-vn));
  }
  return true;
}

__device__ void member_ledger(const KArgs &k, int mi, double time, int flags, int pslot,
                              double (*wsum)[kCh]);

// A member's clock and search phase at the start of a step, from the state
// at the start of the previous step and the flags that step raised (the
// finalize and flush launches advance the member record the same way).
__device__ inline MemberDyn advance_dyn(const MemberDyn &pd, int pf, double dt)
{
  MemberDyn d = pd;
  d.fresh = 0;
  d.pad = 0;
  if (!pd.fresh) {
    d.time = pd.time + dt;
    d.step = pd.step + 1;
    if (pd.mode == MODE_BIN) { d.tab = pd.tab ^ 1; d.mode = MODE_SEARCH; }
    else d.mode = (pf >> 16) ? MODE_BIN : MODE_IDLE;
  }
  return d;
}

// kStage: the first kCandLds candidates of a disk are gathered ahead of the
// bonds and their pair forces staged in LDS (18 KB FP32 / 30 KB FP64 more
// per block, tiles at most kStageMaxTile wide).
template <typename R, bool kDeferred, bool kStage>
__device__ __forceinline__ void plate_step_body(const KArgs &k)
{
  __shared__ double2 sx[kStage ? kMaxLdsStage : kMaxLds];
  __shared__ double ss[kStage ? kMaxRingStage : kMaxRing];
  __shared__ double red[4][kCh];
  __shared__ tdm_plate_member sm;
  __shared__ int sflags[2];
  __shared__ MemberDyn sdyn;
  __shared__ int4 sprevf;
  // per disk: the first kCandLds candidate indices and their signed pair forces
  __shared__ int cQ[kCandLds][kThreads];
  __shared__ R cFx[kCandLds][kThreads], cFy[kCandLds][kThreads];

  const tdm_plate_arrays &a = k.a;
  if (kDeferred && (int)blockIdx.x >= a.nblocks) {
    // one extra block per member carries the member forward (this step's
    // clock and search phase for the next launch) and reduces the previous
    // step's ledger, whose partials and flags are complete now; kept apart
    // from the tiles so that no tile waits on the reduction
    const int mi = (int)blockIdx.x - a.nblocks;
    MemberDyn pd;
    int pf = 0;
    if (k.first) {
      const tdm_plate_member *mp = a.members + mi;
      pd.time = mp->time;
      pd.step = mp->step_count;
      pd.mode = mp->mode;
      pd.tab = mp->tab;
      pd.fresh = 1;
      pd.pad = 0;
    } else {
      pd = k.w.dyn[2 * mi + (k.parity ^ 1)];
      const int4 pf4 = reinterpret_cast<const int4 *>(k.w.mflags)[mi];
      pf = reinterpret_cast<const int *>(&pf4)[pd.step % 3];
    }
    const MemberDyn nd = advance_dyn(pd, pf, a.members[mi].dt);
    if (threadIdx.x == 0) {
      k.w.dyn[2 * mi + k.parity] = nd;
      k.w.mflags[4 * mi + (int)((nd.step + 1) % 3)] = 0;
    }
    if (!pd.fresh) member_ledger(k, mi, pd.time, pf, (int)(pd.step & 1), red);
    return;
  }
  // block table entry: (member, tile x0, tile y0, rank), (nx, ny, first, tile_x)
  const int4 blk = k.w.blocks[2 * blockIdx.x];
  const int4 geo = k.w.blocks[2 * blockIdx.x + 1];
  const int mi = blk.x, tx0 = blk.y, ty0 = blk.z;
  const int nx = geo.x, ny = geo.y, first = geo.z, TX = geo.w, TY = kThreads / TX;
  const int lw = TX + 2 * kHalo, lh = TY + 2 * kHalo;
  const int n = a.n;

  // per-disk inputs are issued before the first barrier so their latency
  // overlaps the halo load
  const int ix = tx0 + (int)threadIdx.x % TX, iy = ty0 + (int)threadIdx.x / TX;
  const bool valid = ix < nx && iy < ny;
  const int p = valid ? first + iy * nx + ix : first;
  double2 vp = make_double2(0.0, 0.0);
  uint32_t links = 0;
  bool fixed = false;
  int nc = 0;
  float2 xb = make_float2(0.0f, 0.0f);
  if (valid) {
    vp = k.v_in[p];
    links = a.links[p];
    fixed = a.fixed[p] != 0;
    nc = k.w.ncand[p];
    xb = k.w.xb[p];
  }
  // the member record is read-only during the step: stage it in LDS so that
  // global stores later in the kernel cannot force re-reads
  static_assert(sizeof(tdm_plate_member) % 8 == 0, "member record must be 8-byte sized");
  if (threadIdx.x < sizeof(tdm_plate_member) / 8)
    reinterpret_cast<uint64_t *>(&sm)[threadIdx.x] =
        reinterpret_cast<const uint64_t *>(a.members + mi)[threadIdx.x];
  if (threadIdx.x < 2) sflags[threadIdx.x] = 0;
  if (kDeferred && !k.first && threadIdx.x >= 64 && threadIdx.x < 68) {
    // the member's state at the start of the previous step and the flags
    // that step raised (written by the previous launch)
    const int t = threadIdx.x - 64;
    if (t < 4)
      reinterpret_cast<uint64_t *>(&sdyn)[t] =
          reinterpret_cast<const uint64_t *>(k.w.dyn + 2 * mi + (k.parity ^ 1))[t];
    if (t == 0) sprevf = reinterpret_cast<const int4 *>(k.w.mflags)[mi];
  }
  const int rw = TX + 2, nring = rw * (TY + 2);
  auto ring_site = [&](int t, int &hx, int &hy) { hx = tx0 - 1 + t % rw; hy = ty0 - 1 + t / rw; };
  // stage positions of tile + 2-disk halo
  // (at most kMaxLds / kThreads + 1 = 6 rounds; loads of a round are issued together)
  for (int t0 = 0; t0 < lw * lh; t0 += 2 * kThreads) {
    double2 v[2];
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const int t = t0 + u * kThreads + threadIdx.x;
      const int hx = tx0 - kHalo + t % lw, hy = ty0 - kHalo + t / lw;
      v[u] = make_double2(0.0, 0.0);
      if (t < lw * lh && on_lattice(hx, hy, nx, ny)) v[u] = k.x_in[first + hy * nx + hx];
    }
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const int t = t0 + u * kThreads + threadIdx.x;
      if (t < lw * lh) sx[t] = v[u];
    }
  }
  __syncthreads();
  const tdm_plate_member &m = sm;
  const double two_r = 2 * m.radius;
  const double half_span = (double)(nx - 1) / 2;
  // clock and search phase of this step: from the member record, or (inside
  // a multi-step call) advanced from the previous step's state exactly as
  // the finalize advances the record
  int mode = m.mode, tab = m.tab;
  double time = m.time;
  long long step = m.step_count;
  if (kDeferred && !k.first) {
    const MemberDyn pd = sdyn;
    const MemberDyn cd =
        advance_dyn(pd, reinterpret_cast<const int *>(&sprevf)[pd.step % 3], m.dt);
    mode = cd.mode; tab = cd.tab; time = cd.time; step = cd.step;
  }
  Tile tile{sx, lw, tx0, ty0, nx, ny, half_span, two_r};
  for (int t = threadIdx.x; t < nring; t += kThreads) {
    int hx, hy;
    ring_site(t, hx, hy);
    // FP64 in both precisions: the strain stage is a small part of the
    // step, and exact strains make every shear-break decision the FP64 one
    double s = 0;
    if (on_lattice(hx, hy, nx, ny)) s = strain64(tile, hx, hy);
    ss[t] = s;
  }
  __syncthreads();

  double ch[kCh];
#pragma unroll
  for (int c = 0; c < kCh; ++c) ch[c] = 0.0;
  // dissipation of unbonded candidate contacts, summed apart from that of
  // broken bonds (list order), so the ledger does not depend on kStage
  double dsc = 0.0;
  int status = 0, req = 0;
  if (valid) {
    const double2 xp = tile.x(ix, iy);
    const double2 sp = lattice_site(ix, iy, half_span, two_r);
    const double s_self = ss[(iy - ty0 + 1) * rw + (ix - tx0 + 1)];
    const uint32_t present = present_links(ix, iy, nx, ny);
    links &= present;
    const double kb = m.bond_stiffness, kc = m.contact_stiffness, cdamp = m.contact_damping;
    const double eps_t_p = m.per_disk_strength ? a.eps_t_disk[p] : m.breaking_strain;
    const double eps_s_p = m.per_disk_strength ? a.eps_s_disk[p] : m.shear_breaking_strain;
    double2 ld = make_double2(0.0, 0.0);
    if (m.has_loads) ld = reinterpret_cast<const double2 *>(a.loads)[p];

    // ---- unbonded contact candidates -------------------------------------
    // Candidate lists live in global memory, slot-major (cand[e*n + p]); the
    // owning thread is the only reader and writer of its slots.
    int *cand = k.w.cand + p;
    const int slots = a.cand_slots;
    if (mode == MODE_SEARCH) {
      nc = 0;
      const double cutoff = two_r + m.skin;
      const double h = m.cell_size;
      int cx = (int)floor((xp.x - m.grid_origin[0]) / h);
      int cy = (int)floor((xp.y - m.grid_origin[1]) / h);
      if (!(xp.x == xp.x) || !(xp.y == xp.y)) { cx = 0; cy = 0; }
      cx = min(max(cx, 0), m.ncx - 1);
      cy = min(max(cy, 0), m.ncy - 1);
      const int *cnt = k.w.cell_count + (size_t)tab * a.ncells;
      const int *ctab = k.w.cell_tab + (size_t)tab * a.ncells * a.cell_cap;
      // the nine cell counts are fetched together; the 3x3 block's entries
      // are then visited kScan at a time (one table load and one position
      // gather per batch) and inserted into a sorted list held in registers;
      // cells o = 0..8 hold entries pre[o]..pre[o+1]-1, entry t of cell o at
      // table offset base[o] + t
      int pre[10], base[9];
      pre[0] = 0;
#pragma unroll
      for (int o = 0; o < 9; ++o) {
        const int gx = cx + o % 3 - 1, gy = cy + o / 3 - 1;
        const bool in = gx >= 0 && gx < m.ncx && gy >= 0 && gy < m.ncy;
        const int cell = in ? m.cell_first + gy * m.ncx + gx : 0;
        pre[o + 1] = in ? min(cnt[cell], a.cell_cap) : 0;
        base[o] = cell * a.cell_cap;
      }
#pragma unroll
      for (int o = 0; o < 9; ++o) {
        base[o] -= pre[o];
        pre[o + 1] += pre[o];
      }
      const int total = pre[9];
      int nl[kMaxCand];
#pragma unroll
      for (int e = 0; e < kMaxCand; ++e) nl[e] = INT_MAX;
      for (int t0 = 0; t0 < total; t0 += kScan) {
        int qs[kScan];
#pragma unroll
        for (int u = 0; u < kScan; ++u) {
          const int t = t0 + u;
          qs[u] = p;
          if (t < total) {
            // the cell holding entry t is the last one starting at or before
            // t (an empty cell starts where the next one does)
            int off = base[0];
#pragma unroll
            for (int c = 1; c < 9; ++c) off = t >= pre[c] ? base[c] : off;
            qs[u] = ctab[(size_t)(off + t)];
          }
        }
        double2 xq[kScan];
#pragma unroll
        for (int u = 0; u < kScan; ++u) xq[u] = qs[u] != p ? k.x_in[qs[u]] : xp;
#pragma unroll
        for (int u = 0; u < kScan; ++u) {
          const int q = qs[u];
          if (q == p) continue;
          const int dq = q - p;
          bool bonded = false;
#pragma unroll
          for (int c = 0; c < 8; ++c)
            if (((present >> c) & 1u) && dq == dyc(c) * nx + dxc(c)) bonded = true;
          if (bonded) continue;
          const double dx = xq[u].x - xp.x, dy = xq[u].y - xp.y;
          if (!(sqrt(dx * dx + dy * dy) < cutoff)) continue;
          if (nc == slots) { status |= TDM_ST_CAND_OVERFLOW; continue; }
          int carry = q;
#pragma unroll
          for (int e = 0; e < kMaxCand; ++e) {
            const int lo = min(nl[e], carry), hi = max(nl[e], carry);
            nl[e] = lo;
            carry = hi;
          }
          ++nc;
        }
      }
#pragma unroll
      for (int e = 0; e < kMaxCand; ++e)
        if (e < nc) cand[(size_t)e * n] = nl[e];
      k.w.ncand[p] = (uint8_t)nc;
      xb = make_float2((float)(xp.x - sp.x), (float)(xp.y - sp.y));
      k.w.xb[p] = xb;
    } else {
      double ddx = (xp.x - sp.x) - (double)xb.x, ddy = (xp.y - sp.y) - (double)xb.y;
      double lim = 0.5 * m.skin * 0.999;
      if (ddx * ddx + ddy * ddy >= lim * lim) status |= TDM_ST_STALE_LIST;
    }
    int split = 0;  // list slots below split hold partners below p
    // unit rest vector lo -> hi of p and a non-neighbour q (the normal of
    // coincident disks): the two lattice sites' difference, dem.py:119-129
    auto rest_dir = [&](int q, bool self_lo) {
      const int ql = q - first, jx = ql % nx, jy = ql / nx;
      const double2 sq = lattice_site(jx, jy, half_span, two_r);
      const double2 s_lo = self_lo ? sp : sq, s_hi = self_lo ? sq : sp;
      const double ix0 = s_hi.x - s_lo.x, iy0 = s_hi.y - s_lo.y;
      const double id = sqrt(ix0 * ix0 + iy0 * iy0);
      return make_double2(ix0 / id, iy0 / id);
    };

    // unbonded candidate q, lo -> hi oriented; accumulates into f
    auto unbonded = [&](int q, double &fx, double &fy, bool energy) {
      const bool self_lo = p < q;
      const double2 xq = k.x_in[q];
      if (energy) {
        if (self_lo) {
          double dx = xq.x - xp.x, dy = xq.y - xp.y;
          double ov = two_r - sqrt(dx * dx + dy * dy);
          if (ov > 0) ch[E_CONTACT] += ov * ov;
        }
        return;
      }
      const double2 vq = k.v_in[q];
      double Fx, Fy, diss;
      if (contact_pair<R>(self_lo ? xp : xq, self_lo ? xq : xp, self_lo ? vp : vq,
                          self_lo ? vq : vp, [&] { return rest_dir(q, self_lo); }, m.radius, kc,
                          cdamp, Fx, Fy, diss)) {
        if (self_lo) { fx += Fx; fy += Fy; dsc += diss; }
        else { fx += -Fx; fy += -Fy; }
      }
    };

    if (k.op == OP_ENERGY) {
      while (split < nc && cand[(size_t)split * n] < p) ++split;
      // dem.py:603-624: each bond and contact counted once, at its lower end
#pragma unroll
      for (int c = 4; c < 8; ++c) {
        if (!((present >> c) & 1u)) continue;
        const double2 xq = tile.x(ix + dxc(c), iy + dyc(c));
        const double2 sq = lattice_site(ix + dxc(c), iy + dyc(c), half_span, two_r);
        double rx = sq.x - sp.x, ry = sq.y - sp.y;
        double rest = sqrt(rx * rx + ry * ry);
        double dx = xq.x - xp.x, dy = xq.y - xp.y;
        double d = sqrt(dx * dx + dy * dy);
        if ((links >> c) & 1u) { double e = d - rest; ch[E_BOND] += e * e; }
        else { double ov = two_r - d; if (ov > 0) ch[E_CONTACT] += ov * ov; }
      }
      double dummy_x = 0, dummy_y = 0;
      for (int e = split; e < nc; ++e) unbonded(cand[(size_t)e * n], dummy_x, dummy_y, true);
      const double ox = m.tool_start[0] + time * m.tool_velocity[0];
      const double oy = m.tool_start[1] + time * m.tool_velocity[1];
      double d;
      if (m.hull_m == 0) {
        double dx = xp.x - ox, dy = xp.y - oy;
        d = sqrt(dx * dx + dy * dy);
      } else {
        const double2 *hv = reinterpret_cast<const double2 *>(a.hull) + m.hull_first;
        d = INFINITY;
        for (int s2 = 0; s2 < m.hull_m - 1; ++s2) {
          double sx0 = hv[s2].x + ox, sy0 = hv[s2].y + oy;
          double gx = (hv[s2 + 1].x + ox) - sx0, gy = (hv[s2 + 1].y + oy) - sy0;
          double fr = ((xp.x - sx0) * gx + (xp.y - sy0) * gy) / (gx * gx + gy * gy);
          fr = fr < 0.0 ? 0.0 : (fr > 1.0 ? 1.0 : fr);
          double ddx = xp.x - (sx0 + fr * gx), ddy = xp.y - (sy0 + fr * gy);
          double dd = sqrt(ddx * ddx + ddy * ddy);
          if (dd < d) d = dd;
        }
      }
      double ov = m.radius + m.tool_radius - d;
      if (ov > 0) ch[E_TOOL] += ov * ov;
      ch[E_VSQ] += vp.x * vp.x + vp.y * vp.y;
    } else {
      // ---- unbonded candidates first: the first kCandLds list slots are
      // gathered kCandBatch at a time (indices, then partner state) and their
      // signed pair forces are left in this thread's LDS slots for the
      // ordered sum below, so a rubble disk waits on a few gather rounds
      // instead of two dependent loads per candidate; slots past kCandLds
      // (rare) are evaluated inline during the sum
      uint32_t chas = 0;
      const int ncl = min(nc, kCandLds);
      if constexpr (kStage) {
        for (int e0 = 0; e0 < ncl; e0 += kCandBatch) {
          int qs[kCandBatch];
  #pragma unroll
          for (int u = 0; u < kCandBatch; ++u) qs[u] = e0 + u < ncl ? cand[(size_t)(e0 + u) * n] : p;
          double2 xq[kCandBatch], vq[kCandBatch];
  #pragma unroll
          for (int u = 0; u < kCandBatch; ++u) {
            xq[u] = xp;
            vq[u] = vp;
            if (e0 + u < ncl) {
              xq[u] = k.x_in[qs[u]];
              vq[u] = k.v_in[qs[u]];
            }
          }
  #pragma unroll
          for (int u = 0; u < kCandBatch; ++u) {
            const int e = e0 + u;
            if (e >= ncl) continue;
            const int q = qs[u];
            cQ[e][threadIdx.x] = q;
            const bool self_lo = p < q;
            double Fx, Fy, diss;
            if (contact_pair<R>(self_lo ? xp : xq[u], self_lo ? xq[u] : xp, self_lo ? vp : vq[u],
                                self_lo ? vq[u] : vp, [&] { return rest_dir(q, self_lo); },
                                m.radius, kc, cdamp, Fx, Fy, diss)) {
              // the signed contribution the sum adds (FP32 forces are float
              // values, so the narrowing store is exact)
              chas |= 1u << e;
              cFx[e][threadIdx.x] = (R)(self_lo ? Fx : -Fx);
              cFy[e][threadIdx.x] = (R)(self_lo ? Fy : -Fy);
              if (self_lo) dsc += diss;
            }
          }
        }
      }
      auto cand_at = [&](int e) -> int {
        if constexpr (kStage) {
          if (e < kCandLds) return cQ[e][threadIdx.x];
        }
        return cand[(size_t)e * n];
      };
      auto cand_add = [&](int e, int qc, double &fx_, double &fy_) {
        if constexpr (kStage) {
          if (e < kCandLds) {
            if ((chas >> e) & 1u) {
              fx_ += (double)cFx[e][threadIdx.x];
              fy_ += (double)cFy[e][threadIdx.x];
            }
            return;
          }
        }
        unbonded(qc, fx_, fy_, false);
      };
      while (split < nc && cand_at(split) < p) ++split;

      // ---- all present bonds, independently (4-way instruction-level
      // parallelism per group); fracture commit and force per bond --------
      // per-bond forces in the force precision (FP32 values widen exactly
      // when they are summed into the FP64 total)
      R Fx[8], Fy[8];
      uint32_t has = 0;
      // one bond: fracture commit and force (c is a constant after unrolling)
      auto bond = [&](const int c) {
        Fx[c] = 0.0;
        Fy[c] = 0.0;
        if (!((present >> c) & 1u)) return;
        const bool self_lo = c >= 4;
        const int jx = ix + dxc(c), jy = iy + dyc(c);
        const int q = p + dyc(c) * nx + dxc(c);
        const double2 xq = tile.x(jx, jy);
        const double s_q = ss[(jy - ty0 + 1) * rw + (jx - tx0 + 1)];
        const double2 x_lo = self_lo ? xp : xq, x_hi = self_lo ? xq : xp;
        const double sh_lo = self_lo ? s_self : s_q, sh_hi = self_lo ? s_q : s_self;
        // exact rest vector from the two lattice sites (dem.py:119-129)
        auto site_vec = [&](double &rx_, double &ry_, double &rest_) {
          const double2 sq = lattice_site(jx, jy, half_span, two_r);
          const double2 s_lo = self_lo ? sp : sq, s_hi = self_lo ? sq : sp;
          rx_ = s_hi.x - s_lo.x;
          ry_ = s_hi.y - s_lo.y;
          rest_ = sqrt(rx_ * rx_ + ry_ * ry_);
        };
        double eps_t = m.breaking_strain, eps_s = m.shear_breaking_strain;
        if (m.per_disk_strength) {
          eps_t = fmin(eps_t_p, a.eps_t_disk[q]);
          eps_s = fmin(eps_s_p, a.eps_s_disk[q]);
        }
        bool alive = (links >> c) & 1u;
        // exact FP64 geometry of the pair (the FP32 path uses it for break
        // decisions near a threshold and for the failure record)
        auto geom64 = [&](double rx, double ry, double rest, double &d, double &nxv,
                          double &nyv) {
          double dx = x_hi.x - x_lo.x, dy = x_hi.y - x_lo.y;
          d = sqrt(dx * dx + dy * dy);
          if (d > 0) { double dd = d > kEps64 ? d : kEps64; nxv = dx / dd; nyv = dy / dd; }
          else { nxv = rx / rest; nyv = ry / rest; }
        };
        auto record = [&](int tf, int sf, double d64, double rest) {
          // the lower-index end owns the bond's failure record (dem.py:259-277)
          links &= ~(1u << c);
          if (!self_lo) return;
          const double ext64 = d64 - rest;
          const int b = a.fwd_bond[p] + __popc(present & ((1u << c) - 1u) & 0xF0u);
          a.alive[b] = 0;
          a.failure_mode[b] = (uint8_t)(tf + 2 * sf);
          a.failure_time_s[b] = time;
          a.failure_extension_m[b] = ext64;
          const double en = (0.5 * kb) * (ext64 * ext64);
          a.failure_energy_J[b] = en;
          ch[C_RELEASED] += en;
          double ov = two_r - d64;
          if (ov < 0) ov = 0;
          ch[C_ACQUIRED] += (0.5 * kc) * (ov * ov);
          ch[C_BREAKS] += 1.0;
        };
        // a bond broken before this step is a contact only: its geometry and
        // break test are skipped (rubble tiles run whole waves of those)
        auto broken_contact = [&]() {
          const double2 vq = k.v_in[q];
          double diss;
          auto n0 = [&] {
            double rx, ry, rest;
            site_vec(rx, ry, rest);
            return make_double2(rx / rest, ry / rest);
          };
          if (contact_pair<R>(x_lo, x_hi, self_lo ? vp : vq, self_lo ? vq : vp, n0, m.radius, kc,
                              cdamp, Fx[c], Fy[c], diss)) {
            has |= 1u << c;
            if (self_lo) ch[C_PAIR_DISS] += diss;
          }
        };
        if constexpr (Prec<R>::fp64) {
          if (alive) {
            double rx, ry, rest, d, nxv, nyv;
            site_vec(rx, ry, rest);
            geom64(rx, ry, rest, d, nxv, nyv);
            double ext = d - rest;
            double tstrain = ext / rest;
            double bshear = 0.5 * (sh_lo + sh_hi);
            int sf = bshear > eps_s, tf = tstrain > eps_t;
            if (sf || tf) {
              alive = false;
              record(tf, sf, d, rest);
            } else {
              double mag = kb * ext;
              Fx[c] = mag * nxv;
              Fy[c] = mag * nyv;
              has |= 1u << c;
            }
          }
          if (!alive) broken_contact();
        } else {
          // FP32: displacement relative to the nominal lattice vector keeps
          // the extension free of cancellation; exact FP64 site geometry is
          // only formed for break decisions inside the margin and records.
          if (alive) {
            const double rnx = (self_lo ? dxc(c) : -dxc(c)) * two_r;
            const double rny = (self_lo ? dyc(c) : -dyc(c)) * two_r;
            const float rfx = (float)rnx, rfy = (float)rny;
            const float restf = (dxc(c) != 0 && dyc(c) != 0) ? (float)(two_r * 1.4142135623730951)
                                                               : (float)two_r;
            float ux = (float)((x_hi.x - x_lo.x) - rnx), uy = (float)((x_hi.y - x_lo.y) - rny);
            float dvx = rfx + ux, dvy = rfy + uy;
            float d = sqrt32(dvx * dvx + dvy * dvy);
            float ext = (2.0f * (rfx * ux + rfy * uy) + (ux * ux + uy * uy)) * rcp32(d + restf);
            const float inv_rest = rcp32(restf);
            float nxv, nyv;
            if (d > 0) { float r = rcp32(fmaxf(d, kEps32)); nxv = dvx * r; nyv = dvy * r; }
            else { nxv = rfx * inv_rest; nyv = rfy * inv_rest; }
            // the shear test is the FP64 one (strains are FP64); the tensile
            // test is redone in FP64 inside the margin of its threshold
            const float tstrain = ext * inv_rest;
            const double mg = m.fp32_margin;
            const bool near_t = tstrain >= (float)(eps_t * (1.0 - mg)) &&
                                tstrain <= (float)(eps_t * (1.0 + mg));
            int tf = tstrain > (float)eps_t;
            const int sf = 0.5 * (sh_lo + sh_hi) > eps_s;
            if (near_t) {
              double rx, ry, rest, d64, n1, n2;
              site_vec(rx, ry, rest);
              geom64(rx, ry, rest, d64, n1, n2);
              tf = ((d64 - rest) / rest) > eps_t;
            }
            if (sf || tf) {
              alive = false;
              double rx, ry, rest, d64, n1, n2;
              site_vec(rx, ry, rest);
              geom64(rx, ry, rest, d64, n1, n2);
              record(tf, sf, d64, rest);
            } else {
              float mag = (float)kb * ext;
              Fx[c] = mag * nxv;
              Fy[c] = mag * nyv;
              has |= 1u << c;
            }
          }
          if (!alive) broken_contact();
        }
      };

      // ---- sum in index_add_ order: -drag*v, then +F of partners above
      // (ascending), then -F of partners below (ascending); unbonded
      // candidates are merged in by partner index --------------------------
      // bonds are evaluated in two groups of four (partners above, then
      // below), each summed before the next is formed: fewer live registers
      double fx = (-m.drag) * vp.x, fy = (-m.drag) * vp.y;
      int ci = split;
#pragma unroll
      for (int c = 4; c < 8; ++c) bond(c);
#pragma unroll
      for (int c = 4; c < 8; ++c) {
        if (!((present >> c) & 1u)) continue;
        const int q = p + dyc(c) * nx + dxc(c);
        while (ci < nc) {
          const int qc = cand_at(ci);
          if (qc > q) break;
          cand_add(ci, qc, fx, fy);
          ++ci;
        }
        if ((has >> c) & 1u) { fx += (double)Fx[c]; fy += (double)Fy[c]; }
      }
      for (; ci < nc; ++ci) cand_add(ci, cand_at(ci), fx, fy);
      ci = 0;
#pragma unroll
      for (int c = 0; c < 4; ++c) bond(c);
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        if (!((present >> c) & 1u)) continue;
        const int q = p + dyc(c) * nx + dxc(c);
        while (ci < split) {
          const int qc = cand_at(ci);
          if (qc > q) break;
          cand_add(ci, qc, fx, fy);
          ++ci;
        }
        if ((has >> c) & 1u) { fx += -(double)Fx[c]; fy += -(double)Fy[c]; }
      }
      for (; ci < split; ++ci) cand_add(ci, cand_at(ci), fx, fy);
      a.links[p] = (uint8_t)links;

      // ---- tool or hull (dem.py:294-324, hull_contact.py:7-75) ------------
      const double tvx = m.tool_velocity[0], tvy = m.tool_velocity[1];
      double tfx, tfy, cfx, cfy;
      if (m.hull_m == 0) {
        const double tpx = m.tool_start[0] + time * m.tool_velocity[0];
        const double tpy = m.tool_start[1] + time * m.tool_velocity[1];
        double dx = xp.x - tpx, dy = xp.y - tpy;
        if constexpr (Prec<R>::fp64) {
          double d = sqrt(dx * dx + dy * dy);
          double nxv, nyv;
          if (d > 0) { double dd = d > kEps64 ? d : kEps64; nxv = dx / dd; nyv = dy / dd; }
          else { nxv = 0.0; nyv = -1.0; }
          double ov = m.radius + m.tool_radius - d;
          if (ov < 0) ov = 0;
          double sp2 = (vp.x - tvx) * nxv + (vp.y - tvy) * nyv;
          double mag = kc * ov - cdamp * sp2;
          if (mag < 0) mag = 0;
          if (!(ov > 0)) mag = 0;
          tfx = mag * nxv; tfy = mag * nyv;
          cfx = kc * ov * nxv; cfy = kc * ov * nyv;
        } else {
          float dxf = (float)dx, dyf = (float)dy;
          float d = sqrt32(dxf * dxf + dyf * dyf);
          float nxv, nyv;
          if (d > 0) { float r = rcp32(fmaxf(d, kEps32)); nxv = dxf * r; nyv = dyf * r; }
          else { nxv = 0.0f; nyv = -1.0f; }
          float ov = fmaxf((float)(m.radius + m.tool_radius) - d, 0.0f);
          float sp2 = (float)(vp.x - tvx) * nxv + (float)(vp.y - tvy) * nyv;
          float kcf = (float)kc;
          float mag = ov > 0 ? fmaxf(kcf * ov - (float)cdamp * sp2, 0.0f) : 0.0f;
          tfx = (double)(mag * nxv); tfy = (double)(mag * nyv);
          cfx = (double)(kcf * ov * nxv); cfy = (double)(kcf * ov * nyv);
        }
      } else {
        const double ox = m.tool_start[0] + time * m.tool_velocity[0];
        const double oy = m.tool_start[1] + time * m.tool_velocity[1];
        const double2 *hv = reinterpret_cast<const double2 *>(a.hull) + m.hull_first;
        int best = 0;
        double best_d2 = INFINITY, bdx = 0, bdy = 0;
        for (int s2 = 0; s2 < m.hull_m - 1; ++s2) {
          double sx0 = hv[s2].x + ox, sy0 = hv[s2].y + oy;
          double gx = (hv[s2 + 1].x + ox) - sx0, gy = (hv[s2 + 1].y + oy) - sy0;
          double len2 = gx * gx + gy * gy;
          double fr = ((xp.x - sx0) * gx + (xp.y - sy0) * gy) / len2;
          fr = fr < 0.0 ? 0.0 : (fr > 1.0 ? 1.0 : fr);
          double ddx = xp.x - (sx0 + fr * gx), ddy = xp.y - (sy0 + fr * gy);
          double d2 = ddx * ddx + ddy * ddy;
          if (d2 < best_d2) { best_d2 = d2; best = s2; bdx = ddx; bdy = ddy; }
        }
        double distance = sqrt(best_d2);
        double gx = (hv[best + 1].x + ox) - (hv[best].x + ox);
        double gy = (hv[best + 1].y + oy) - (hv[best].y + oy);
        double len = sqrt(gx * gx + gy * gy);
        double ttx = gx / len, tty = gy / len;
        if (ttx < 0 || (ttx == 0 && tty < 0)) { ttx = -ttx; tty = -tty; }
        double fbx = -tty, fby = ttx;
        double rvx = vp.x - tvx, rvy = vp.y - tvy;
        if (rvx * fbx + rvy * fby > 0) { fbx = -fbx; fby = -fby; }
        double nxv, nyv;
        if (distance > kEps64) { nxv = bdx / distance; nyv = bdy / distance; }
        else { nxv = fbx; nyv = fby; }
        double ov = m.radius + m.tool_radius - distance;
        if (ov < 0) ov = 0;
        double sp2 = rvx * nxv + rvy * nyv;
        double mag = kc * ov - cdamp * sp2;
        if (mag < 0) mag = 0;
        if (!(ov > 0)) mag = 0;
        tfx = mag * nxv; tfy = mag * nyv;
        double cm = kc * ov;  // damping 0; the clamp is a no-op for ov >= 0
        if (!(ov > 0)) cm = 0;
        cfx = cm * nxv; cfy = cm * nyv;
      }
      ch[C_TOOL_FX] += tfx; ch[C_TOOL_FY] += tfy;
      ch[C_TOOL_DISS] += (tfx - cfx) * (vp.x - tvx) + (tfy - cfy) * (vp.y - tvy);
      ch[C_VSQ] += vp.x * vp.x + vp.y * vp.y;
      ch[C_VSUM_X] += vp.x; ch[C_VSUM_Y] += vp.y;
      fx += tfx; fy += tfy;
      if (m.has_loads) {
        fx += ld.x; fy += ld.y;
        ch[C_LOAD_X] += ld.x; ch[C_LOAD_Y] += ld.y;
        ch[C_MOMENT] += xp.x * ld.y - xp.y * ld.x;
      }
      if (fixed) { ch[C_SUP_X] += fx; ch[C_SUP_Y] += fy; }

      if (k.op == OP_STEP) {
        // dem.py:363-366
        double vx, vy;
        if constexpr (Prec<R>::fp64) {
          vx = vp.x + (m.dt * fx) / m.mass;  // operand order of dem.py:363
          vy = vp.y + (m.dt * fy) / m.mass;
        } else {
          const double dt_m = m.dt / m.mass;
          vx = vp.x + fx * dt_m;
          vy = vp.y + fy * dt_m;
        }
        if (fixed) { vx = 0.0; vy = 0.0; }
        double xn = xp.x + m.dt * vx, yn = xp.y + m.dt * vy;
        if (fixed) { xn = sp.x; yn = sp.y; }
        k.x_out[p] = make_double2(xn, yn);
        k.v_out[p] = make_double2(vx, vy);
        if (m.has_loads) ch[C_WORK] += ld.x * (xn - xp.x) + ld.y * (yn - xp.y);
        if (!isfinite(xn) || !isfinite(yn) || !isfinite(vx) || !isfinite(vy)) status |= TDM_ST_NONFINITE;
        // rebuild request: quarter skin since the candidate list was built
        double ddx = (xn - sp.x) - (double)xb.x, ddy = (yn - sp.y) - (double)xb.y;
        double q4 = 0.25 * m.skin;
        if (ddx * ddx + ddy * ddy >= q4 * q4) req = 1;
        if (mode == MODE_BIN) {
          const double h = m.cell_size;
          int cx = (int)floor((xn - m.grid_origin[0]) / h);
          int cy = (int)floor((yn - m.grid_origin[1]) / h);
          if (!(xn == xn) || !(yn == yn)) { cx = 0; cy = 0; }
          cx = min(max(cx, 0), m.ncx - 1);
          cy = min(max(cy, 0), m.ncy - 1);
          int cell = m.cell_first + cy * m.ncx + cx;
          int t = tab ^ 1;
          int slot = atomicAdd(k.w.cell_count + (size_t)t * a.ncells + cell, 1);
          if (slot < a.cell_cap) k.w.cell_tab[((size_t)t * a.ncells + cell) * a.cell_cap + slot] = p;
          else status |= TDM_ST_CELL_OVERFLOW;
        }
      } else {
        reinterpret_cast<double2 *>(a.force)[p] = make_double2(fx, fy);
      }
    }
  }

  // a search step clears the other hash table for the next bin phase
  if (mode == MODE_SEARCH) {
    const int ncell_m = m.ncx * m.ncy;
    int *cnt = k.w.cell_count + (size_t)(tab ^ 1) * a.ncells + m.cell_first;
    for (int c = blk.w * kThreads + threadIdx.x; c < ncell_m; c += m.nblocks * kThreads) cnt[c] = 0;
  }
  if (status) atomicOr(&sflags[0], status);
  if (req) atomicOr(&sflags[1], 1);
  ch[C_PAIR_DISS] += dsc;
  block_sum(ch, red);
  const int pslot = kDeferred ? (int)(step & 1) : 0;
  if (threadIdx.x == 0) {
    const size_t slot = (size_t)(m.block_first + blk.w);
#pragma unroll
    for (int c = 0; c < kCh; ++c)
      k.w.partials[((size_t)pslot * kCh + c) * a.nblocks + slot] = ch[c];
    const int f = sflags[0] | (sflags[1] << 16);
    if (!kDeferred) k.w.bflags[slot] = f;
    else if (f) atomicOr(k.w.mflags + 4 * mi + (int)(step % 3), f);
  }
}

// 4 waves per SIMD: both precisions fit 128 VGPRs without spills (the
// FP32 build keeps its FP64 fallback to the tensile test only, and strains
// are FP64 in both), so a grid of several rounds of blocks hides latency.
template <typename R, bool kDeferred, bool kStage>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(kStage ? 2 : 4)))
void plate_step_kernel(KArgs k)
{
  plate_step_body<R, kDeferred, kStage>(k);
}

// ---------------------------------------------------------------------------
// Ledger reductions.  A member's step ledger is reduced by one 256-thread
// block: each thread sums its blocks' 16 channels, then a fixed DPP tree per
// wave and the 4 wave totals in order, so the order depends only on the
// member's block count (independent of the batch around it, and the same
// whether the reduction runs in the finalize launch of a single step, in the
// next step's ledger block or in the flush at the end of a multi-step call).
// Separate launches make every block's partials visible without device-scope
// fences inside a step (such fences write back the L2 and cost more than a
// launch).
// ---------------------------------------------------------------------------
constexpr int kFinThreads = kThreads;

__device__ void reduce_partials(const KArgs &k, int b0, int nb, int pslot, double (*wsum)[kCh],
                                double *stot)
{
  double sum[kCh];
#pragma unroll
  for (int c = 0; c < kCh; ++c) sum[c] = 0.0;
  for (int b = threadIdx.x; b < nb; b += kFinThreads) {
    double v[kCh];
#pragma unroll
    for (int c = 0; c < kCh; ++c)
      v[c] = k.w.partials[((size_t)pslot * kCh + c) * k.a.nblocks + b0 + b];
#pragma unroll
    for (int c = 0; c < kCh; ++c) sum[c] += v[c];
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int c = 0; c < kCh; ++c) {
    double t = 0.0;
    if (__ballot(sum[c] != 0.0)) t = wave_sum(sum[c]);
    if (lane == 0) wsum[wave][c] = t;
  }
  __syncthreads();
  if (threadIdx.x < kCh) {
    double t = 0.0;
#pragma unroll
    for (int w = 0; w < kFinThreads / 64; ++w) t += wsum[w][threadIdx.x];
    stot[threadIdx.x] = t;
  }
  __syncthreads();
}

// Step ledger of member mi (dem.py:362-378): impulses, work, damping, the
// last-step record and the status bits; `time` is the step's start time.
// Called by every thread of a block; thread 0 writes.
__device__ void member_ledger(const KArgs &k, int mi, double time, int flags, int pslot,
                              double (*wsum)[kCh])
{
  __shared__ double stot[kCh];
  const tdm_plate_arrays &a = k.a;
  tdm_plate_member *mp = a.members + mi;
  const int b0 = mp->block_first, nb = mp->nblocks;
  __syncthreads();  // wsum may be the caller's reduction scratch
  reduce_partials(k, b0, nb, pslot, wsum, stot);
  if (threadIdx.x != 0) return;
  const double dt = mp->dt, drag = mp->drag;
  const int has_loads = mp->has_loads;
  const double tvx = mp->tool_velocity[0], tvy = mp->tool_velocity[1];
  double *accp = a.accounting + (size_t)mi * TDM_NACC;
  double acc[TDM_NACC];
#pragma unroll
  for (int c = 0; c < TDM_NACC; ++c) acc[c] = accp[c];
  double tot[kCh];
#pragma unroll
  for (int q = 0; q < kCh; ++q) tot[q] = stot[q];
  acc[TDM_ACC_RELEASE] += tot[C_RELEASED];
  acc[TDM_ACC_ACQUIRED] += tot[C_ACQUIRED];
  const double rx_ = -tot[C_TOOL_FX], ry_ = -tot[C_TOOL_FY];
  const double supx = -tot[C_SUP_X], supy = -tot[C_SUP_Y];
  const double extx = has_loads ? tot[C_LOAD_X] : 0.0, exty = has_loads ? tot[C_LOAD_Y] : 0.0;
  const double dimx = (-drag) * tot[C_VSUM_X], dimy = (-drag) * tot[C_VSUM_Y];
  acc[TDM_ACC_EXT_IMP] = acc[TDM_ACC_EXT_IMP] + dt * extx;
  acc[TDM_ACC_EXT_IMP + 1] = acc[TDM_ACC_EXT_IMP + 1] + dt * exty;
  acc[TDM_ACC_SUP_IMP] = acc[TDM_ACC_SUP_IMP] + dt * supx;
  acc[TDM_ACC_SUP_IMP + 1] = acc[TDM_ACC_SUP_IMP + 1] + dt * supy;
  acc[TDM_ACC_DRAG_IMP] = acc[TDM_ACC_DRAG_IMP] + dt * dimx;
  acc[TDM_ACC_DRAG_IMP + 1] = acc[TDM_ACC_DRAG_IMP + 1] + dt * dimy;
  acc[TDM_ACC_TOOL_IMP] = acc[TDM_ACC_TOOL_IMP] + dt * (-rx_);
  acc[TDM_ACC_TOOL_IMP + 1] = acc[TDM_ACC_TOOL_IMP + 1] + dt * (-ry_);
  const double damp = (tot[C_PAIR_DISS] + (-tot[C_TOOL_DISS])) + drag * tot[C_VSQ];
  const double tpow = -(rx_ * tvx + ry_ * tvy);
  const double work = has_loads ? tot[C_WORK] : 0.0;
  acc[TDM_ACC_EXT_WORK] += work;
  acc[TDM_ACC_TOOL_WORK] += dt * tpow;
  acc[TDM_ACC_DAMPING] += dt * damp;
  double last[TDM_NLAST] = {time, time + dt, work, supx, supy, rx_, ry_, extx, exty,
                            has_loads ? tot[C_MOMENT] : 0.0};
  double *lastp = a.last + (size_t)mi * TDM_NLAST;
#pragma unroll
  for (int c = 0; c < TDM_NLAST; ++c) lastp[c] = last[c];
#pragma unroll
  for (int c = 0; c < TDM_NACC; ++c) accp[c] = acc[c];
  if (flags & 0xFFFF) atomicOr(&mp->status, flags & 0xFFFF);
}

// clock and search phase after a step that started in (time, step, mode, tab)
// and raised `flags`
__device__ inline void advance_record(tdm_plate_member *mp, double time, long long step, int mode,
                                      int tab, int flags)
{
  mp->time = time + mp->dt;
  mp->step_count = step + 1;
  if (mode == MODE_BIN) { mp->tab = tab ^ 1; mp->mode = MODE_SEARCH; }
  else { mp->tab = tab; mp->mode = (flags >> 16) ? MODE_BIN : MODE_IDLE; }
}

// Finalize of a single launch (one step, a force evaluation or an energy
// evaluation): one block per member.
__global__ __launch_bounds__(kFinThreads) void plate_finalize_kernel(KArgs k)
{
  __shared__ int sflag;
  __shared__ double wsum[kFinThreads / 64][kCh];
  __shared__ double stot[kCh];
  const tdm_plate_arrays &a = k.a;
  const int mi = blockIdx.x;
  tdm_plate_member *mp = a.members + mi;
  const int b0 = mp->block_first, nb = mp->nblocks;
  if (threadIdx.x == 0) sflag = 0;
  __syncthreads();
  int fl = 0;
  for (int b = threadIdx.x; b < nb; b += kFinThreads) fl |= k.w.bflags[b0 + b];
  if (fl) atomicOr(&sflag, fl);
  __syncthreads();
  const int flags = sflag;
  if (k.op == OP_STEP) {
    const double time = mp->time;
    const long long step = mp->step_count;
    const int mode = mp->mode, tab = mp->tab;
    member_ledger(k, mi, time, flags, 0, wsum);
    if (threadIdx.x == 0) advance_record(mp, time, step, mode, tab, flags);
    return;
  }
  reduce_partials(k, b0, nb, 0, wsum, stot);
  if (threadIdx.x != 0) return;
  const tdm_plate_member m = *mp;
  double tot[kCh];
#pragma unroll
  for (int q = 0; q < kCh; ++q) tot[q] = stot[q];
  if (k.op == OP_ENERGY) {
    double *en = a.energy + (size_t)mi * 5;
    double e0 = 0.5 * m.mass * tot[E_VSQ];
    double e1 = 0.5 * m.bond_stiffness * tot[E_BOND];
    double e2 = 0.5 * m.contact_stiffness * tot[E_CONTACT];
    double e3 = 0.5 * m.contact_stiffness * tot[E_TOOL];
    en[0] = e0; en[1] = e1; en[2] = e2; en[3] = e3;
    en[4] = ((e0 + e1) + e2) + e3;
  } else {
    double *accp = a.accounting + (size_t)mi * TDM_NACC;
    accp[TDM_ACC_RELEASE] += tot[C_RELEASED];
    accp[TDM_ACC_ACQUIRED] += tot[C_ACQUIRED];
    a.reaction[2 * mi] = -tot[C_TOOL_FX];
    a.reaction[2 * mi + 1] = -tot[C_TOOL_FY];
  }
  // the launch searched for candidates if the member was in the search phase
  if (m.mode == MODE_SEARCH) mp->mode = MODE_IDLE;
  if (flags & 0xFFFF) mp->status = m.status | (flags & 0xFFFF);
}

// End of a multi-step call: ledger of the last step, the member record
// brought to the state after it, and the member's flag words cleared for the
// next call (its first launch raises flags into a word it does not clear).
// k.parity is the last launch's parity.
__global__ __launch_bounds__(kFinThreads) void plate_flush_kernel(KArgs k)
{
  __shared__ double wsum[kFinThreads / 64][kCh];
  const int mi = blockIdx.x;
  const MemberDyn st = k.w.dyn[2 * mi + k.parity];
  const int f = k.w.mflags[4 * mi + (int)(st.step % 3)];
  member_ledger(k, mi, st.time, f, (int)(st.step & 1), wsum);
  if (threadIdx.x == 0) {
    advance_record(k.a.members + mi, st.time, st.step, st.mode, st.tab, f);
    reinterpret_cast<int4 *>(k.w.mflags)[mi] = make_int4(0, 0, 0, 0);
  }
}

// ---------------------------------------------------------------------------
// setup helpers
// ---------------------------------------------------------------------------
__global__ void sync_links_kernel(tdm_plate_arrays a, Workspace w)
{
  // one thread per disk: link bit c is alive[bond(p, c)] for every present code
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.n) return;
  // locate the member by binary search over `first`
  int lo = 0, hi = a.nmembers - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (a.members[mid].first <= t) lo = mid; else hi = mid - 1;
  }
  const tdm_plate_member &m = a.members[lo];
  const int l = t - m.first, nx = m.nx, ix = l % nx, iy = l / nx;
  const uint32_t present = present_links(ix, iy, nx, m.ny);
  uint32_t links = 0;
  for (int c = 0; c < 8; ++c) {
    if (!((present >> c) & 1u)) continue;
    int owner = c >= 4 ? t : t + dyc(c) * nx + dxc(c);
    int oc = c >= 4 ? c : 7 - c;
    int ox = owner - m.first;
    uint32_t opres = present_links(ox % nx, ox / nx, nx, m.ny);
    int b = a.fwd_bond[owner] + __popc(opres & ((1u << oc) - 1u) & 0xF0u);
    if (a.alive[b]) links |= 1u << c;
  }
  a.links[t] = (uint8_t)links;
}

__global__ void bin_kernel(tdm_plate_arrays a, Workspace w, const double2 *x)
{
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.n) return;
  int lo = 0, hi = a.nmembers - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (a.members[mid].first <= t) lo = mid; else hi = mid - 1;
  }
  const tdm_plate_member &m = a.members[lo];
  const double2 p = x[t];
  int cx = (int)floor((p.x - m.grid_origin[0]) / m.cell_size);
  int cy = (int)floor((p.y - m.grid_origin[1]) / m.cell_size);
  if (!(p.x == p.x) || !(p.y == p.y)) { cx = 0; cy = 0; }
  cx = min(max(cx, 0), m.ncx - 1);
  cy = min(max(cy, 0), m.ncy - 1);
  int cell = m.cell_first + cy * m.ncx + cx;
  int tb = m.tab;
  int slot = atomicAdd(w.cell_count + (size_t)tb * a.ncells + cell, 1);
  if (slot < a.cell_cap) w.cell_tab[((size_t)tb * a.ncells + cell) * a.cell_cap + slot] = t;
  else atomicOr(&a.members[lo].status, TDM_ST_CELL_OVERFLOW);
}


__global__ void set_mode_kernel(tdm_plate_member *m, int nm, int mode)
{
  int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < nm) m[t].mode = mode;
}

int check_arrays(const tdm_plate_arrays *a, Workspace *w)
{
  if (!a || !a->members || !a->x || !a->v || !a->x_alt || !a->v_alt || !a->links ||
      !a->fixed || !a->fwd_bond || !a->alive || !a->accounting || !a->last || !a->workspace)
    return fail(TDM_EINVAL, "null array in tdm_plate_arrays");
  if (a->n <= 0 || a->nmembers <= 0 || a->nblocks <= 0 || a->ncells <= 0)
    return fail(TDM_EINVAL, "empty batch");
  if (a->cand_slots < 0 || a->cand_slots > kMaxCand || a->cell_cap <= 0)
    return fail(TDM_EINVAL, "cand_slots must be in [0,16] and cell_cap > 0");
  if ((long long)a->ncells * a->cell_cap >= (1LL << 31))
    return fail(TDM_EINVAL, "cell table too large (ncells x cell_cap must stay below 2^31)");
  size_t need = 0;
  carve(a->workspace, a->n, a->nblocks, a->nmembers, a->ncells, a->cand_slots, a->cell_cap, w,
        &need);
  if (need > a->workspace_bytes) return fail(TDM_ESPACE, "workspace too small");
  if (reinterpret_cast<uintptr_t>(a->workspace) & 255)
    return fail(TDM_EINVAL, "workspace must be 256-byte aligned");
  return TDM_OK;
}

// Compute units of the current device (cached per device; filled at setup,
// before any graph capture).
int cu_count()
{
  static std::atomic<int> cached[64];
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
  int c = cached[dev].load(std::memory_order_relaxed);
  if (c <= 0) {
    if (hipDeviceGetAttribute(&c, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        c <= 0)
      c = 256;
    cached[dev].store(c, std::memory_order_relaxed);
  }
  return c;
}

template <typename R>
void (*pick_step(int deferred, bool stage))(KArgs)
{
  if (deferred) return stage ? plate_step_kernel<R, true, true> : plate_step_kernel<R, true, false>;
  return stage ? plate_step_kernel<R, false, true> : plate_step_kernel<R, false, false>;
}

}  // namespace


extern "C" {

int tdm_version(void) { return 100; }

int tdm_last_error(char *buf, size_t len)
{
  if (!buf || len == 0) return TDM_EINVAL;
  std::snprintf(buf, len, "%s", tdm::last_error().c_str());
  return TDM_OK;
}

int tdm_abi_sizes(size_t *member_bytes, size_t *arrays_bytes)
{
  if (!member_bytes || !arrays_bytes) return fail(TDM_EINVAL, "null output");
  *member_bytes = sizeof(tdm_plate_member);
  *arrays_bytes = sizeof(tdm_plate_arrays);
  return TDM_OK;
}

int tdm_plate_tiles(int nx, int ny, int tile_x)
{
  if (nx < 2 || ny < 2 || tile_x < 16 || tile_x > 256 || (kThreads % tile_x))
    return fail(TDM_EINVAL, "tile_x must divide 256 and lie in [16,256]");
  int ty = kThreads / tile_x;
  return ((nx + tile_x - 1) / tile_x) * ((ny + ty - 1) / ty);
}

size_t tdm_plate_workspace_bytes(int n, int nblocks, int nmembers, int ncells, int cand_slots,
                                 int cell_cap)
{
  Workspace w;
  size_t total = 0;
  carve(nullptr, n, nblocks, nmembers, ncells, cand_slots, cell_cap, &w, &total);
  return total;
}

int tdm_plate_setup(const tdm_plate_arrays *a, void *stream)
{
  Workspace w;
  int rc = check_arrays(a, &w);
  if (rc) return rc;
  hipStream_t s = static_cast<hipStream_t>(stream);
  (void)cu_count();  // cached before any launch can be captured into a graph
  // block table: host copy of the member table gives tile counts
  const int nm = a->nmembers;
  tdm_plate_member *hm = new tdm_plate_member[nm];
  hipError_t e = hipMemcpyAsync(hm, a->members, sizeof(tdm_plate_member) * nm,
                                hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) { delete[] hm; return fail(TDM_EHIP, hipGetErrorString(e)); }
  int total = 0;
  for (int i = 0; i < nm; ++i) {
    int t = tdm_plate_tiles(hm[i].nx, hm[i].ny, hm[i].tile_x);
    if (t < 0) { delete[] hm; return t; }
    if (hm[i].block_first != total || hm[i].nblocks != t) {
      delete[] hm;
      return fail(TDM_EINVAL, "member block_first/nblocks do not match tdm_plate_tiles");
    }
    total += t;
  }
  if (total != a->nblocks) { delete[] hm; return fail(TDM_EINVAL, "nblocks mismatch"); }
  int widest = 0;
  for (int i = 0; i < nm; ++i) widest = max(widest, (int)hm[i].tile_x);
  if (a->max_tile_x != 0 && a->max_tile_x < widest) {
    delete[] hm;
    return fail(TDM_EINVAL, "max_tile_x is narrower than a member's tile");
  }
  // Logical tile order keeps a member's tiles contiguous (row-major tiles) so
  // its ledger partials sum in a fixed order; the kernels index partials and
  // flags by logical tile, so the physical launch order is free.  Blocks are
  // dealt round-robin over the 8 XCDs (launch slot s runs on XCD s % 8), and
  // a tile's 2-disk halo is read from its neighbours' rows, so neighbouring
  // tiles should share one XCD's L2: a member of at most kXcdWhole tiles runs
  // entirely on one XCD (the one with the fewest tiles so far), a larger
  // member is cut into 8 contiguous runs of tiles, one per XCD.
  int4 *hb = new int4[2 * total];
  int4 *logical = new int4[total];
  int li = 0;
  for (int i = 0; i < nm; ++i) {
    int tx = hm[i].tile_x, ty = kThreads / tx;
    int ntx = (hm[i].nx + tx - 1) / tx, nty = (hm[i].ny + ty - 1) / ty;
    for (int by = 0; by < nty; ++by)
      for (int bx = 0; bx < ntx; ++bx)
        logical[li++] = make_int4(i, bx * tx, by * ty, by * ntx + bx);
  }
  constexpr int kXcd = 8, kXcdWhole = 32;
  std::vector<int> queue[kXcd];
  for (int i = 0; i < nm; ++i) {
    const int b0 = hm[i].block_first, nb = hm[i].nblocks;
    if (nb <= kXcdWhole) {
      int x = 0;
      for (int c = 1; c < kXcd; ++c)
        if (queue[c].size() < queue[x].size()) x = c;
      for (int b = 0; b < nb; ++b) queue[x].push_back(b0 + b);
    } else {
      for (int r = 0; r < kXcd; ++r)
        for (int b = r * nb / kXcd; b < (r + 1) * nb / kXcd; ++b) queue[r].push_back(b0 + b);
    }
  }
  size_t head[kXcd] = {};
  for (int slot = 0; slot < total; ++slot) {
    int x = slot % kXcd;
    if (head[x] == queue[x].size()) {
      // this XCD's share is spent (shares differ by at most one member):
      // take the next tile of the XCD with the most left
      for (int c = 0; c < kXcd; ++c)
        if (queue[c].size() - head[c] > queue[x].size() - head[x]) x = c;
    }
    const int item = queue[x][head[x]++];
    const int i = logical[item].x;
    hb[2 * slot] = logical[item];
    hb[2 * slot + 1] = make_int4(hm[i].nx, hm[i].ny, hm[i].first, hm[i].tile_x);
  }
  delete[] logical;
  e = hipMemcpyAsync(w.blocks, hb, 2 * sizeof(int4) * total, hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  delete[] hb;
  delete[] hm;
  if (e != hipSuccess) return fail(TDM_EHIP, hipGetErrorString(e));
  // counters, requests, hash tables, candidate counts
  fill_kernel<int><<<64, 256, 0, s>>>(w.bflags, (size_t)a->nblocks, 0);
  fill_kernel<int><<<64, 256, 0, s>>>(w.mflags, 4 * (size_t)a->nmembers, 0);
  fill_kernel<int><<<1024, 256, 0, s>>>(w.cell_count, 2 * (size_t)a->ncells, 0);
  fill_kernel<uint8_t><<<1024, 256, 0, s>>>(w.ncand, (size_t)a->n, (uint8_t)0);
  HIP_TRY(hipGetLastError());
  rc = tdm_plate_sync_links(a, stream);
  if (rc) return rc;
  return tdm_plate_rebin(a, stream);
}

int tdm_plate_sync_links(const tdm_plate_arrays *a, void *stream)
{
  Workspace w;
  int rc = check_arrays(a, &w);
  if (rc) return rc;
  hipStream_t s = static_cast<hipStream_t>(stream);
  sync_links_kernel<<<(a->n + 255) / 256, 256, 0, s>>>(*a, w);
  HIP_TRY(hipGetLastError());
  return TDM_OK;
}

int tdm_plate_rebin(const tdm_plate_arrays *a, void *stream)
{
  // bins the state held in x (parity 0) into each member's current table
  Workspace w;
  int rc = check_arrays(a, &w);
  if (rc) return rc;
  hipStream_t s = static_cast<hipStream_t>(stream);
  fill_kernel<int><<<1024, 256, 0, s>>>(w.cell_count, 2 * (size_t)a->ncells, 0);
  bin_kernel<<<(a->n + 255) / 256, 256, 0, s>>>(*a, w, reinterpret_cast<const double2 *>(a->x));
  set_mode_kernel<<<(a->nmembers + 255) / 256, 256, 0, s>>>(a->members, a->nmembers, MODE_SEARCH);
  HIP_TRY(hipGetLastError());
  return TDM_OK;
}

static KArgs make_args(const tdm_plate_arrays *a, const Workspace &w, int parity, int op,
                       int deferred, int first = 0)
{
  KArgs k;
  k.a = *a;
  k.w = w;
  k.op = op;
  k.parity = parity;
  k.deferred = deferred;
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

// One launch of the step kernel; a single launch (deferred = 0) is followed
// by its finalize.  Candidate forces are staged in LDS when every tile is at
// most kStageMaxTile wide and the staging costs no occupancy: FP32 blocks
// keep four per compute unit; FP64 blocks (three per unit) only when the
// whole grid is resident at two blocks per unit.
static int launch_step(const tdm_plate_arrays *a, const Workspace &w, int parity, int fp32,
                       int op, hipStream_t s, hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr,
                       int deferred = 0, int first = 0)
{
  const KArgs k = make_args(a, w, parity, op, deferred, first);
  const bool stage = a->max_tile_x > 0 && a->max_tile_x <= kStageMaxTile &&
                     (fp32 || a->nblocks <= 2 * cu_count());
  void (*kern)(KArgs) = fp32 ? pick_step<float>(deferred, stage) : pick_step<double>(deferred, stage);
  // a multi-step launch carries one ledger block per member after the tiles
  const int grid = a->nblocks + (deferred ? a->nmembers : 0);
  if (ev0) {
    // events carried by the dispatch itself: they time exactly this kernel
    hipExtLaunchKernelGGL(kern, dim3(grid), dim3(kThreads), 0, s, ev0, ev1, 0, k);
  } else {
    kern<<<grid, kThreads, 0, s>>>(k);
  }
  HIP_TRY(hipGetLastError());
  if (!deferred) {
    plate_finalize_kernel<<<a->nmembers, kFinThreads, 0, s>>>(k);
    HIP_TRY(hipGetLastError());
  }
  return TDM_OK;
}


// nsteps steps from `parity`: one launch per step.  For two or more steps
// the per-step finalize launches are replaced by the deferred ledger (each
// step's ledger blocks reduce the previous step), closed by a flush launch;
// the first launch reads the member records.  ev (optional) holds 2 events
// per step around each step launch.
static int launch_steps(const tdm_plate_arrays *a, const Workspace &w, int nsteps, int parity,
                        int fp32, hipStream_t s, hipEvent_t *ev = nullptr)
{
  if (nsteps <= 0) return TDM_OK;
  if (nsteps == 1) {
    for (int i = 0; i < nsteps; ++i) {
      int rc = launch_step(a, w, parity, fp32, OP_STEP, s, ev ? ev[2 * i] : nullptr,
                           ev ? ev[2 * i + 1] : nullptr);
      if (rc) return rc;
      parity ^= 1;
    }
    return TDM_OK;
  }
  for (int i = 0; i < nsteps; ++i) {
    int rc = launch_step(a, w, parity, fp32, OP_STEP, s, ev ? ev[2 * i] : nullptr,
                         ev ? ev[2 * i + 1] : nullptr, 1, i == 0);
    if (rc) return rc;
    parity ^= 1;
  }
  const KArgs k = make_args(a, w, parity ^ 1, OP_STEP, 1);
  plate_flush_kernel<<<a->nmembers, kFinThreads, 0, s>>>(k);
  HIP_TRY(hipGetLastError());
  return TDM_OK;
}

int tdm_plate_step(const tdm_plate_arrays *a, int nsteps, int parity, int fp32, void *stream)
{
  Workspace w;
  int rc = check_arrays(a, &w);
  if (rc) return rc;
  if (nsteps < 0 || (parity != 0 && parity != 1)) return fail(TDM_EINVAL, "bad nsteps/parity");
  return launch_steps(a, w, nsteps, parity, fp32, static_cast<hipStream_t>(stream));
}

int tdm_plate_step_timed(const tdm_plate_arrays *a, int nsteps, int fp32, void *stream,
                         double *mean_kernel_us)
{
  Workspace w;
  int rc = check_arrays(a, &w);
  if (rc) return rc;
  if (nsteps <= 0 || !mean_kernel_us) return fail(TDM_EINVAL, "bad timing request");
  hipStream_t s = static_cast<hipStream_t>(stream);
  hipEvent_t *ev = new hipEvent_t[2 * nsteps]();
  int made = 0;
  for (; made < 2 * nsteps; ++made)
    if (hipEventCreate(&ev[made]) != hipSuccess) break;
  if (made < 2 * nsteps) rc = fail(TDM_EHIP, "hipEventCreate failed");
  if (rc == TDM_OK) rc = launch_steps(a, w, nsteps, 0, fp32, s, ev);
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

int tdm_plate_forces(const tdm_plate_arrays *a, int parity, int fp32, void *stream)
{
  Workspace w;
  int rc = check_arrays(a, &w);
  if (rc) return rc;
  if (!a->force || !a->reaction) return fail(TDM_EINVAL, "force/reaction outputs required");
  return launch_step(a, w, parity, fp32, OP_FORCES, static_cast<hipStream_t>(stream));
}

int tdm_plate_energy(const tdm_plate_arrays *a, int parity, void *stream)
{
  Workspace w;
  int rc = check_arrays(a, &w);
  if (rc) return rc;
  if (!a->energy) return fail(TDM_EINVAL, "energy output required");
  return launch_step(a, w, parity, 0, OP_ENERGY, static_cast<hipStream_t>(stream));
}

int tdm_plate_graph_create(const tdm_plate_arrays *a, int pairs, int fp32, void *stream,
                           tdm_graph **out)
{
  Workspace w;
  int rc = check_arrays(a, &w);
  if (rc) return rc;
  if (!out || pairs <= 0) return fail(TDM_EINVAL, "bad graph request");
  hipStream_t s = static_cast<hipStream_t>(stream);
  HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  rc = launch_steps(a, w, 2 * pairs, 0, fp32, s);
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

int tdm_graph_launch(tdm_graph *g, void *stream)
{
  if (!g) return fail(TDM_EINVAL, "null graph");
  HIP_TRY(hipGraphLaunch(g->exec, static_cast<hipStream_t>(stream)));
  return TDM_OK;
}

int tdm_graph_destroy(tdm_graph *g)
{
  if (!g) return TDM_OK;
  (void)hipGraphExecDestroy(g->exec);
  (void)hipGraphDestroy(g->graph);
  delete g;
  return TDM_OK;
}

}  // extern "C"
