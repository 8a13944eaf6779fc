// Bonded-sphere lattice in 3-D: one force evaluation (and semi-implicit
// Euler step) of the reference's IceDEM3D (src/tensordem/dem3d.py:1300-1660)
// on the device, FP64, one thread per sphere.
//
// Four launches per evaluation:
//
//   d3_strain_kernel   objective shear strain of every sphere (the
//                      least-squares deformation gradient over its bonds,
//                      dem3d.py:875-901) and the sphere's entry in the hashed
//                      cell table of this evaluation (cell width 2r, so every
//                      overlapping pair lies in 27 adjacent cells);
//   d3_bond_kernel     per bond: extension, fracture commit and failure
//                      record, the bond force on the bond's lower sphere;
//   d3_step_kernel     per sphere: its bond forces gathered in order,
//                      contacts of the 27 cells sorted by partner,
//                      Cundall-Strack springs, tool, platens, integration;
//                      clears the other cell table;
//   d3_finish_kernel   sums the per-block partials in block order and
//                      advances the clock and the lateral platens.
//
// No float atomics: each bond's force is computed once, like the oracle's,
// and a sphere's force is summed in one fixed order (drag,
// gravity, load, bonds to higher partners ascending, bonds to lower partners
// ascending, contacts by ascending partner, tool, platens), the order of the
// oracle, so FP64 runs compare bit for bit.  A pair's spring state is kept
// by both spheres, each computing it from the same operands (lower index
// first), so neither writes the other's data.
#include "common.h"

namespace tdm {
namespace {

constexpr int kSlots = 26;   // lattice neighbours of a sphere
constexpr int kCand = 24;    // candidate partners per sphere (list capacity)
constexpr int kNch = 12;     // partial channels

enum { OP_STEP = 0, OP_FORCES = 1, OP_ENERGY = 2 };
enum { DAMP_NONE = 0, DAMP_RESTITUTION = 1, DAMP_ABSOLUTE = 2 };
enum {
  C_RX = 0, C_RY, C_RZ, C_TOP, C_BOT, C_LAT0, C_LAT1, C_LAT2, C_LAT3, C_FPOW, C_CONTACTS,
  C_BREAKS  // written by the finish kernel from the bond kernel's count
};

struct D3Ws {
  double *inv;        // [n][9]
  uint8_t *degen;     // [n]
  double *shear;      // [n]
  int *count;         // [2][ncells]: cleared by the evaluation before the one that fills it
  int *tab;           // [ncells][cap], filled and read by rebuild evaluations only
  int *hq;            // [2][n][kh]
  double *hxi;        // [2][n][kh][3]
  int *hcnt;          // [2][n]
  double *partials;   // [nblocks][kNch]
  int *status;        // [0] status bits, [1] bonds broken by this evaluation,
                      // [2] rebuild the candidate lists in this evaluation,
                      // [3] rebuild requested by this evaluation, [4] rebuilds
  double *bf;         // [nbonds][3] bond force on the lower sphere
  int *list;          // [n][kCand] partners within 2r + skin at the last build, ascending
  int *lcnt;          // [n]
  double *xb;         // [n][3] positions at the last build
};

__host__ __device__ inline void carve_d3(void *base, int n, int nbonds, int ncells, int cap,
                                         int kh, int nblocks, D3Ws *w, size_t *total)
{
  size_t off = 0;
  char *b = static_cast<char *>(base);
  auto take = [&](size_t bytes) {
    char *q = b ? b + off : nullptr;
    off += align256(bytes);
    return q;
  };
  const size_t N = size_t(n);
  w->inv = reinterpret_cast<double *>(take(sizeof(double) * 9 * N));
  w->degen = reinterpret_cast<uint8_t *>(take(N));
  w->shear = reinterpret_cast<double *>(take(sizeof(double) * N));
  w->count = reinterpret_cast<int *>(take(sizeof(int) * 2 * size_t(ncells)));
  w->tab = reinterpret_cast<int *>(take(sizeof(int) * size_t(ncells) * size_t(cap)));
  w->hq = reinterpret_cast<int *>(take(sizeof(int) * 2 * N * size_t(kh)));
  w->hxi = reinterpret_cast<double *>(take(sizeof(double) * 6 * N * size_t(kh)));
  w->hcnt = reinterpret_cast<int *>(take(sizeof(int) * 2 * N));
  w->partials = reinterpret_cast<double *>(take(sizeof(double) * kNch * size_t(nblocks)));
  w->status = reinterpret_cast<int *>(take(sizeof(int) * 64));
  w->bf = reinterpret_cast<double *>(take(sizeof(double) * 3 * size_t(nbonds > 0 ? nbonds : 1)));
  w->list = reinterpret_cast<int *>(take(sizeof(int) * kCand * N));
  w->lcnt = reinterpret_cast<int *>(take(sizeof(int) * N));
  w->xb = reinterpret_cast<double *>(take(sizeof(double) * 3 * N));
  *total = off;
}

// blocks of one XCD take a contiguous run of spheres (launch slot b runs on
// XCD b % 8), so a sphere's lattice neighbours sit in the same L2
__device__ inline int xcd_block(int b, int nb)
{
  const int x = b & 7, k = b >> 3, q = nb >> 3, r = nb & 7;
  return x * q + (x < r ? x : r) + k;
}

__device__ inline double3 ld3(const double *p, int i)
{
  return make_double3(p[3 * i], p[3 * i + 1], p[3 * i + 2]);
}

__device__ inline void st3(double *p, int i, double3 v)
{
  p[3 * i] = v.x;
  p[3 * i + 1] = v.y;
  p[3 * i + 2] = v.z;
}

__device__ inline double norm3(double3 d) { return sqrt(d.x * d.x + d.y * d.y + d.z * d.z); }

__device__ inline long long cell_coord(double x, double w)
{
  double c = floor(x / w);
  c = fmin(fmax(c, -1.0e15), 1.0e15);
  return (long long)c;
}

__device__ inline int cell_hash(long long cx, long long cy, long long cz, int mask)
{
  unsigned long long h = (unsigned long long)cx * 73856093ull ^ (unsigned long long)cy * 19349663ull
                         ^ (unsigned long long)cz * 83492791ull;
  return int(h & (unsigned long long)mask);
}

// ---------------------------------------------------------------- setup
__global__ void d3_covinv_kernel(tdm_d3_arrays a, tdm_d3_params p, D3Ws w)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= p.n) return;
  // sum r r^T over the bonds to higher partners, then lower partners
  // (dem3d.py:883-885 adds over the bond's i, then its j, in bond order)
  double c[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  const double3 xi = ld3(a.x0, i);
  for (int s = 0; s < kSlots; ++s) {
    const int q = a.nbr[size_t(i) * kSlots + s];
    if (q < 0) {
      if (s < 13) { s = 12; continue; }
      break;
    }
    const double3 xq = ld3(a.x0, q);
    const double3 r = s < 13 ? make_double3(xq.x - xi.x, xq.y - xi.y, xq.z - xi.z)
                             : make_double3(xi.x - xq.x, xi.y - xq.y, xi.z - xq.z);
    const double rr[3] = {r.x, r.y, r.z};
#pragma unroll
    for (int u = 0; u < 3; ++u)
#pragma unroll
      for (int v = 0; v < 3; ++v) c[3 * u + v] += rr[u] * rr[v];
  }
  const double A0 = c[0], B0 = c[1], C0 = c[2], D0 = c[4], E0 = c[5], F0 = c[8];
  const double A = D0 * F0 - E0 * E0, B = C0 * E0 - B0 * F0, C = B0 * E0 - C0 * D0;
  const double D = A0 * F0 - C0 * C0, E = B0 * C0 - A0 * E0, F = A0 * D0 - B0 * B0;
  const double det = A0 * A + B0 * B + C0 * C;
  const double tr = A0 + D0 + F0;
  double *o = w.inv + size_t(i) * 9;
  if (!(det > 1e-10 * tr * tr * tr)) {
    w.degen[i] = 1;
    for (int k = 0; k < 9; ++k) o[k] = 0.0;
    return;
  }
  w.degen[i] = 0;
  o[0] = A / det; o[1] = B / det; o[2] = C / det;
  o[3] = B / det; o[4] = D / det; o[5] = E / det;
  o[6] = C / det; o[7] = E / det; o[8] = F / det;
}

// ---------------------------------------------------------------- strain + cell table
__global__ __launch_bounds__(kThreads) void d3_strain_kernel(tdm_d3_arrays a, tdm_d3_params p,
                                                             D3Ws w, const double *x, int tab)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= p.n) return;
  const double3 xi = ld3(x, i);
  if (tab < 0 || !w.status[2]) {
    // strain only: no candidate search in this evaluation
  } else if (!isfinite(xi.x) || !isfinite(xi.y) || !isfinite(xi.z)) {
    atomicOr(w.status, TDM_ST_NONFINITE);
  } else {
    const int h = cell_hash(cell_coord(xi.x, p.cell_w), cell_coord(xi.y, p.cell_w),
                            cell_coord(xi.z, p.cell_w), p.ncells - 1);
    const int slot = atomicAdd(w.count + size_t(tab) * p.ncells + h, 1);
    if (slot < p.cell_cap)
      w.tab[size_t(h) * p.cell_cap + slot] = i;
    else
      atomicOr(w.status, TDM_ST_CELL_OVERFLOW);
  }
  // objective shear strain (dem3d.py:875-901): sum q r^T over the bonds to
  // higher partners, then lower partners, each in ascending partner order;
  // absent slots (-1, after the present ones) load the sphere itself and add
  // nothing, so every gather of a half is issued at once
  double cr[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  const double3 x0i = ld3(a.x0, i);
  const int *row = a.nbr + size_t(i) * kSlots;
#pragma unroll
  for (int half = 0; half < 2; ++half) {
    int qs[13];
#pragma unroll
    for (int k = 0; k < 13; ++k) qs[k] = row[half * 13 + k];
#pragma unroll
    for (int k = 0; k < 13; ++k) {
      const bool present = qs[k] >= 0;
      const int q = present ? qs[k] : i;
      const double3 xq = ld3(x, q), x0q = ld3(a.x0, q);
      double qq[3], rr[3];
      if (half == 0) {
        qq[0] = xq.x - xi.x; qq[1] = xq.y - xi.y; qq[2] = xq.z - xi.z;
        rr[0] = x0q.x - x0i.x; rr[1] = x0q.y - x0i.y; rr[2] = x0q.z - x0i.z;
      } else {
        qq[0] = xi.x - xq.x; qq[1] = xi.y - xq.y; qq[2] = xi.z - xq.z;
        rr[0] = x0i.x - x0q.x; rr[1] = x0i.y - x0q.y; rr[2] = x0i.z - x0q.z;
      }
      if (present) {
#pragma unroll
        for (int u = 0; u < 3; ++u)
#pragma unroll
          for (int v = 0; v < 3; ++v) cr[3 * u + v] += qq[u] * rr[v];
      }
    }
  }
  if (w.degen[i]) {
    w.shear[i] = 0.0;
    return;
  }
  const double *m = w.inv + size_t(i) * 9;
  double mi[9];
#pragma unroll
  for (int k = 0; k < 9; ++k) mi[k] = m[k];
  double F[9], e[9];
#pragma unroll
  for (int u = 0; u < 3; ++u)
#pragma unroll
    for (int v = 0; v < 3; ++v)
      F[3 * u + v] = cr[3 * u] * mi[v] + cr[3 * u + 1] * mi[3 + v] + cr[3 * u + 2] * mi[6 + v];
#pragma unroll
  for (int u = 0; u < 3; ++u)
#pragma unroll
    for (int v = 0; v < 3; ++v) {
      const double ftf = F[u] * F[v] + F[3 + u] * F[3 + v] + F[6 + u] * F[6 + v];
      e[3 * u + v] = 0.5 * (ftf - (u == v ? 1.0 : 0.0));
    }
  const double mean = (e[0] + e[4] + e[8]) / 3.0;
  double s2 = 0.0;
#pragma unroll
  for (int u = 0; u < 3; ++u)
#pragma unroll
    for (int v = 0; v < 3; ++v) {
      const double d = e[3 * u + v] - (u == v ? mean : 0.0);
      s2 += d * d;
    }
  w.shear[i] = sqrt((2.0 / 3.0) * s2);
}

// ---------------------------------------------------------------- bonds
// One thread per bond: extension, the fracture commit (dem3d.py:984-1009;
// the bond's own record, one writer) and the bond force on its lower sphere
// (dem3d.py:1327-1350), which both spheres then gather.
__global__ __launch_bounds__(kThreads) void d3_bond_kernel(tdm_d3_arrays a, tdm_d3_params p, D3Ws w,
                                                           const double *x, int commit)
{
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= p.nbonds) return;
  const int i = a.pairs[2 * b], j = a.pairs[2 * b + 1];
  const double3 xi = ld3(x, i), xj = ld3(x, j), x0i = ld3(a.x0, i), x0j = ld3(a.x0, j);
  const double3 d = make_double3(xj.x - xi.x, xj.y - xi.y, xj.z - xi.z);
  const double3 d0 = make_double3(x0j.x - x0i.x, x0j.y - x0i.y, x0j.z - x0i.z);
  const double dist = norm3(d), l0 = norm3(d0);
  const double ext = dist - l0;
  if (commit && a.alive[b]) {
    const double thr = a.thr ? a.thr[b] : p.eps_t;
    const bool tensile = ext / l0 > thr;
    const bool shear = 0.5 * (w.shear[i] + w.shear[j]) > p.eps_s;
    if (tensile || shear) {
      a.alive[b] = 0;
      a.failure_mode[b] = uint8_t(int(tensile) + 2 * int(shear));
      a.failure_time_s[b] = a.clock[0];
      a.failure_extension_m[b] = ext;
      a.failure_energy_J[b] = 0.5 * p.kb * (ext * ext);
      atomicAdd(w.status + 1, 1);
    }
  }
  const double mag = p.kb * ext;
  const double den = l0 > kEps64 ? l0 : kEps64;
  const double nx = dist > kEps64 ? d.x / dist : d0.x / den;
  const double ny = dist > kEps64 ? d.y / dist : d0.y / den;
  const double nz = dist > kEps64 ? d.z / dist : d0.z / den;
  w.bf[3 * size_t(b)] = mag * nx;
  w.bf[3 * size_t(b) + 1] = mag * ny;
  w.bf[3 * size_t(b) + 2] = mag * nz;
}

// ---------------------------------------------------------------- step
struct StepArgs {
  const double *x, *v;
  double *x_out, *v_out;
  const int *hq_in, *hcnt_in;
  const double *hxi_in;
  int *hq_out, *hcnt_out;
  double *hxi_out;
  int tab, op, commit, upd_hist;
};

__global__ __launch_bounds__(kThreads) void d3_step_kernel(tdm_d3_arrays a, tdm_d3_params p,
                                                           D3Ws w, StepArgs s)
{
  __shared__ int cand[kCand][kThreads];
  __shared__ double red[4][kCh];
  const int t = threadIdx.x;
  const int blk = xcd_block(blockIdx.x, gridDim.x);
  const int i = blk * kThreads + t;
  const bool live = i < p.n;
  double ch[kNch];
#pragma unroll
  for (int k = 0; k < kNch; ++k) ch[k] = 0.0;

  // clear the other cell table for the next evaluation
  {
    int *other = w.count + size_t(s.tab ^ 1) * p.ncells;
    for (int c = blockIdx.x * kThreads + t; c < p.ncells; c += gridDim.x * kThreads) other[c] = 0;
  }

  const double time = a.clock[0];
  if (live) {
    const double r = p.radius, two_r = 2.0 * r;
    const double3 xi = ld3(s.x, i), vi = ld3(s.v, i);
    const bool fixed_i = a.fixed[i] != 0;
    double f[3] = {-p.drag * vi.x, -p.drag * vi.y, -p.drag * vi.z};
    if (p.gravity) f[2] += -p.mg;
    if (a.loads) {
      f[0] += a.loads[3 * i];
      f[1] += a.loads[3 * i + 1];
      f[2] += a.loads[3 * i + 2];
    }
    const double3 x0i = ld3(a.x0, i);
    const int *row = a.nbr + size_t(i) * kSlots;
    const int *brow = a.bond_of_slot + size_t(i) * kSlots;
    // bond forces of the bond kernel: bonds to higher partners added, to
    // lower partners subtracted, each in ascending partner order
    uint32_t alive_mask = 0;
#pragma unroll
    for (int half = 0; half < 2; ++half) {
      int bs[13];
#pragma unroll
      for (int k = 0; k < 13; ++k) bs[k] = brow[half * 13 + k];
#pragma unroll
      for (int k = 0; k < 13; ++k) {
        const bool present = bs[k] >= 0;
        const size_t b = present ? size_t(bs[k]) : 0;
        const bool alive = present && a.alive[b] != 0;
        const double bx = w.bf[3 * b], by = w.bf[3 * b + 1], bz = w.bf[3 * b + 2];
        if (alive) {
          alive_mask |= 1u << (half * 13 + k);
          if (half == 0) {
            f[0] += bx; f[1] += by; f[2] += bz;
          } else {
            f[0] -= bx; f[1] -= by; f[2] -= bz;
          }
        }
      }
    }

    // candidate partners: every sphere within 2r + skin when the lists were
    // last built, ascending.  A rebuild searches the 27 cells (of width
    // 2r + skin) around the sphere; it is due when a sphere has moved 0.45
    // skin since the last one, so every pair closer than 2r is listed.
    const bool rebuild = w.status[2] != 0;
    const double reach = two_r + p.skin;
    int *lrow = w.list + size_t(i) * kCand;
    int nc = 0;
    if (rebuild) {
      const long long cx = cell_coord(xi.x, p.cell_w), cy = cell_coord(xi.y, p.cell_w),
                      cz = cell_coord(xi.z, p.cell_w);
      const int *cnt = w.count + size_t(s.tab) * p.ncells;
      const int *tb = w.tab;
      for (int c = 0; c < 27; ++c) {
        // two of the 27 cells may share a bucket: a sphere found twice is
        // dropped when it is inserted
        const int h = cell_hash(cx + (c % 3) - 1, cy + ((c / 3) % 3) - 1, cz + (c / 9) - 1,
                                p.ncells - 1);
        const int m = min(cnt[h], p.cell_cap);
        for (int e = 0; e < m; ++e) {
          const int q = tb[size_t(h) * p.cell_cap + e];
          if (q == i) continue;
          const double3 xq = ld3(s.x, q);
          const double3 d = make_double3(xq.x - xi.x, xq.y - xi.y, xq.z - xi.z);
          if (fabs(d.x) >= reach || fabs(d.y) >= reach || fabs(d.z) >= reach) continue;
          if (!(norm3(d) < reach)) continue;
          bool seen = false;
          for (int e = 0; e < nc; ++e) seen |= cand[e][t] == q;
          if (seen) continue;
          if (nc >= kCand) {
            atomicOr(w.status, TDM_ST_CAND_OVERFLOW);
            continue;
          }
          int pos = nc++;
          while (pos > 0 && cand[pos - 1][t] > q) {
            cand[pos][t] = cand[pos - 1][t];
            --pos;
          }
          cand[pos][t] = q;
        }
      }
      for (int e = 0; e < nc; ++e) lrow[e] = cand[e][t];
      w.lcnt[i] = nc;
      st3(w.xb, i, xi);
    } else {
      nc = w.lcnt[i];
    }

    const int kh = p.kh;
    int hin_n = 0, hin_pos = 0, hout_n = 0;
    if (p.friction) hin_n = s.hcnt_in[i];
    double epair = 0.0;
    for (int c = 0; c < nc; ++c) {
      const int q = lrow[c];
      const bool lower = i < q;
      const int ia = lower ? i : q, ja = lower ? q : i;
      const double3 xa = ld3(s.x, ia), xb = ld3(s.x, ja);
      const double3 cd = make_double3(xb.x - xa.x, xb.y - xa.y, xb.z - xa.z);
      if (fabs(cd.x) >= two_r || fabs(cd.y) >= two_r || fabs(cd.z) >= two_r) continue;
      const double cdist = norm3(cd);
      const double ov = two_r - cdist;
      if (!(ov > 0.0)) continue;
      // pairs whose bond is alive are not contacts (dem3d.py:1374-1378)
      bool bonded = false;
      for (int sl = 0; sl < kSlots; ++sl)
        if (row[sl] == q) bonded = (alive_mask >> sl) & 1u;
      if (bonded) continue;
      if (s.op == OP_ENERGY) {
        if (lower) epair += p.hertz ? p.hz_pair_u * (ov * ov * sqrt(ov)) : (0.5 * p.kc) * (ov * ov);
        continue;
      }
      const double den = cdist > kEps64 ? cdist : kEps64;
      const double3 cn = make_double3(cd.x / den, cd.y / den, cd.z / den);
      const double3 va = ld3(s.v, ia), vb = ld3(s.v, ja);
      const double3 rv = make_double3(vb.x - va.x, vb.y - va.y, vb.z - va.z);
      const double vn = rv.x * cn.x + rv.y * cn.y + rv.z * cn.z;
      const bool pinned = (a.fixed[ia] != 0) != (a.fixed[ja] != 0);
      const double meff = pinned ? p.mass : p.mass / 2.0;
      const double kloc = p.hertz ? p.hz_pair * sqrt(ov) : p.kc;
      double gam = 0.0;
      if (p.damp == DAMP_ABSOLUTE) gam = p.cdamp;
      else if (p.damp == DAMP_RESTITUTION)
        gam = p.hertz ? p.zeta2 * sqrt(kloc * meff) : (pinned ? p.gamma_pinned : p.gamma_mobile);
      double mag = kloc * ov - gam * vn;
      if (mag < 0.0) mag = 0.0;
      double ft[3] = {0.0, 0.0, 0.0};
      if (p.friction) {
        // Cundall-Strack spring (dem3d.py:1420-1493); the stored spring of the
        // pair is found by merging the partner-sorted lists
        double td[3] = {0.0, 0.0, 0.0};
        while (hin_pos < hin_n && s.hq_in[size_t(i) * kh + hin_pos] < q) ++hin_pos;
        if (hin_pos < hin_n && s.hq_in[size_t(i) * kh + hin_pos] == q) {
          const double *o = s.hxi_in + (size_t(i) * kh + hin_pos) * 3;
          td[0] = o[0]; td[1] = o[1]; td[2] = o[2];
        }
        const double vt[3] = {rv.x - vn * cn.x, rv.y - vn * cn.y, rv.z - vn * cn.z};
        if (s.upd_hist) {
          td[0] = td[0] + vt[0] * p.dt;
          td[1] = td[1] + vt[1] * p.dt;
          td[2] = td[2] + vt[2] * p.dt;
        }
        const double kt = p.mindlin_kt ? p.hz_kt * sqrt(ov) : p.kt;
        const double raw[3] = {kt * td[0], kt * td[1], kt * td[2]};
        const double fm = sqrt(raw[0] * raw[0] + raw[1] * raw[1] + raw[2] * raw[2]);
        const double lim = p.mu * mag;
        const double scale = fm > lim ? lim / (fm > 1e-15 ? fm : 1e-15) : 1.0;
        ft[0] = raw[0] * scale; ft[1] = raw[1] * scale; ft[2] = raw[2] * scale;
        if (lower)
          ch[C_FPOW] += ft[0] * (va.x - vb.x) + ft[1] * (va.y - vb.y) + ft[2] * (va.z - vb.z);
        if (s.upd_hist) {
          if (hout_n >= kh) {
            atomicOr(w.status, TDM_ST_HIST_OVERFLOW);
          } else {
            s.hq_out[size_t(i) * kh + hout_n] = q;
            double *o = s.hxi_out + (size_t(i) * kh + hout_n) * 3;
            o[0] = ft[0] / kt; o[1] = ft[1] / kt; o[2] = ft[2] / kt;
            ++hout_n;
          }
        }
      }
      if (lower) ch[C_CONTACTS] += 1.0;
      const double cf[3] = {-mag * cn.x, -mag * cn.y, -mag * cn.z};
      if (lower) {
        f[0] += ft[0]; f[1] += ft[1]; f[2] += ft[2];
        f[0] += cf[0]; f[1] += cf[1]; f[2] += cf[2];
      } else {
        f[0] -= ft[0]; f[1] -= ft[1]; f[2] -= ft[2];
        f[0] -= cf[0]; f[1] -= cf[1]; f[2] -= cf[2];
      }
    }
    if (s.upd_hist) s.hcnt_out[i] = hout_n;

    if (s.op == OP_ENERGY) {
      a.pair_energy[i] = epair;
    } else {
      // tool (dem3d.py:1529-1543)
      {
        const double *tl = a.tool;
        const double td[3] = {xi.x - (tl[0] + time * tl[3]), xi.y - (tl[1] + time * tl[4]),
                              xi.z - (tl[2] + time * tl[5])};
        const double tdist = sqrt(td[0] * td[0] + td[1] * td[1] + td[2] * td[2]);
        const double den = tdist > kEps64 ? tdist : kEps64;
        double tn[3] = {td[0] / den, td[1] / den, td[2] / den};
        if (!(tdist > kEps64)) { tn[0] = 0.0; tn[1] = 0.0; tn[2] = -1.0; }
        double tov = r + p.tool_r - tdist;
        if (tov < 0.0) tov = 0.0;
        const double tvn = (vi.x - tl[3]) * tn[0] + (vi.y - tl[4]) * tn[1] + (vi.z - tl[5]) * tn[2];
        const double el = p.hertz ? p.hz_rigid * (tov * sqrt(tov)) : p.kc * tov;
        double tm = el - p.cdamp * tvn;
        if (tm < 0.0) tm = 0.0;
        if (!(tov > 0.0)) tm = 0.0;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          const double tf = tm * tn[k];
          ch[C_RX + k] -= tf;
          f[k] += tf;
        }
      }
      // top and bottom platens (dem3d.py:1545-1593)
      if (p.top) {
        const double z = p.top_z0 + time * p.top_v;
        double ov = xi.z + r - z;
        if (ov < 0.0) ov = 0.0;
        const double el = p.hertz ? p.hz_rigid * (ov * sqrt(ov)) : p.platen_k * ov;
        double mg = el + p.platen_c * (vi.z - p.top_v);
        if (mg < 0.0) mg = 0.0;
        if (!(ov > 0.0)) mg = 0.0;
        f[2] += -mg;
        ch[C_TOP] += mg;
      }
      if (p.bot) {
        const double z = p.bot_z0 + time * p.bot_v;
        double ov = z + r - xi.z;
        if (ov < 0.0) ov = 0.0;
        const double el = p.hertz ? p.hz_rigid * (ov * sqrt(ov)) : p.platen_k * ov;
        double mg = el - p.platen_c * (vi.z - p.bot_v);
        if (mg < 0.0) mg = 0.0;
        if (!(ov > 0.0)) mg = 0.0;
        f[2] += mg;
        ch[C_BOT] -= mg;
      }
      // lateral platens x_min, x_max, y_min, y_max (dem3d.py:1595-1637)
      if (p.lateral) {
        const double *lp = a.clock + 2, *lv = a.clock + 6;
#pragma unroll
        for (int wl = 0; wl < 4; ++wl) {
          const int ax = wl < 2 ? 0 : 1;
          const double xc = ax == 0 ? xi.x : xi.y;
          const double vc = (ax == 0 ? vi.x : vi.y) - lv[ax == 0 ? 0 : 2];
          double ov = (wl & 1) ? xc + r - lp[wl] : lp[wl] - xc + r;
          if (ov < 0.0) ov = 0.0;
          double mg = (wl & 1) ? p.lat_k * ov + p.lat_c * vc : p.lat_k * ov - p.lat_c * vc;
          if (mg < 0.0) mg = 0.0;
          if (!(ov > 0.0)) mg = 0.0;
          if (wl & 1) {
            f[ax] -= mg;
            ch[C_LAT0 + wl] += mg;
          } else {
            f[ax] += mg;
            ch[C_LAT0 + wl] -= mg;
          }
        }
      }
      st3(a.force, i, make_double3(f[0], f[1], f[2]));
      if (s.op == OP_STEP) {
        // dem3d.py:1647-1653
        double xn[3], vnw[3];
        const double xo[3] = {xi.x, xi.y, xi.z}, vo[3] = {vi.x, vi.y, vi.z};
        const double x0[3] = {x0i.x, x0i.y, x0i.z};
        bool finite = true;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
          double acc = f[k] / p.mass;
          if (fixed_i) acc = 0.0;
          double vv = vo[k] + acc * p.dt;
          if (fixed_i) vv = 0.0;
          double xx = xo[k] + vv * p.dt;
          if (fixed_i) xx = x0[k];
          vnw[k] = vv;
          xn[k] = xx;
          finite &= isfinite(xx) && isfinite(vv);
        }
        if (!finite) atomicOr(w.status, TDM_ST_NONFINITE);
        const double3 x_b = ld3(w.xb, i);
        const double dx = xn[0] - x_b.x, dy = xn[1] - x_b.y, dz = xn[2] - x_b.z;
        const double lim = 0.45 * p.skin;
        if (!(dx * dx + dy * dy + dz * dz <= lim * lim)) atomicOr(w.status + 3, 1);
        st3(s.x_out, i, make_double3(xn[0], xn[1], xn[2]));
        st3(s.v_out, i, make_double3(vnw[0], vnw[1], vnw[2]));
      }
    }
  }
  if (s.op == OP_ENERGY) return;
  block_sum<kNch>(ch, red);
  if (t == 0)
    for (int k = 0; k < kNch; ++k) w.partials[size_t(blk) * kNch + k] = ch[k];
}

// ---------------------------------------------------------------- finish
__global__ __launch_bounds__(kThreads) void d3_finish_kernel(tdm_d3_arrays a, tdm_d3_params p,
                                                             D3Ws w, int op, int nblocks)
{
  __shared__ double red[4][kCh];
  double ch[kNch];
#pragma unroll
  for (int k = 0; k < kNch; ++k) ch[k] = 0.0;
  for (int b = threadIdx.x; b < nblocks; b += kThreads)
#pragma unroll
    for (int k = 0; k < kNch; ++k) ch[k] += w.partials[size_t(b) * kNch + k];
  block_sum<kNch>(ch, red);
  if (threadIdx.x != 0) return;
  for (int k = 0; k < kNch; ++k) a.out[k] = ch[k];
  a.out[C_BREAKS] = double(w.status[1]);
  w.status[1] = 0;
  if (w.status[2]) w.status[4] += 1;
  w.status[2] = w.status[3];
  w.status[3] = 0;
  a.out[TDM_D3_STATUS] = double(*w.status);
  a.out[TDM_D3_REBUILDS] = double(w.status[4]);
  if (op == OP_STEP) {
    double *c = a.clock;
    if (p.lateral)
      for (int k = 0; k < 4; ++k) c[2 + k] = c[2 + k] + c[6 + k] * p.dt;
    c[0] = c[0] + p.dt;
    c[1] = c[1] + 1.0;
  }
}

__global__ void d3_copy_kernel(double *dst, const double *src, size_t n)
{
  size_t t = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
  for (; t < n; t += size_t(gridDim.x) * blockDim.x) dst[t] = src[t];
}

__global__ void d3_status_kernel(const D3Ws w, double *out)
{
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  if (w.status[2]) w.status[4] += 1;
  w.status[2] = w.status[3];
  w.status[3] = 0;
  out[TDM_D3_STATUS] = double(*w.status);
  out[TDM_D3_REBUILDS] = double(w.status[4]);
}

int check_d3(const tdm_d3_arrays *a, const tdm_d3_params *p, D3Ws *w)
{
  if (!a || !p) return fail(TDM_EINVAL, "null arrays or params");
  if (p->n < 1 || p->nbonds < 0) return fail(TDM_EINVAL, "bad sphere or bond count");
  if (p->ncells < 1 || (p->ncells & (p->ncells - 1))) return fail(TDM_EINVAL, "ncells must be a power of two");
  if (p->cell_cap < 1 || p->kh < 1) return fail(TDM_EINVAL, "bad cell or history capacity");
  if (p->nblocks != tdm_d3_blocks(p->n)) return fail(TDM_EINVAL, "nblocks does not match n");
  if (!(p->skin > 0.0) || !(p->cell_w >= 2.0 * p->radius + p->skin))
    return fail(TDM_EINVAL, "cell width must cover 2 r + skin, skin positive");
  if (!a->x || !a->v || !a->x_alt || !a->v_alt || !a->x0 || !a->fixed || !a->nbr ||
      !a->bond_of_slot || !a->force || !a->clock || !a->out || !a->tool || !a->workspace)
    return fail(TDM_EINVAL, "missing array");
  if (p->nbonds > 0 && (!a->alive || !a->pairs || !a->failure_mode || !a->failure_time_s ||
                        !a->failure_extension_m || !a->failure_energy_J))
    return fail(TDM_EINVAL, "missing bond array");
  size_t need = 0;
  carve_d3(nullptr, p->n, p->nbonds, p->ncells, p->cell_cap, p->kh, p->nblocks, w, &need);
  if (a->workspace_bytes < need) return fail(TDM_ESPACE, "workspace too small");
  if (reinterpret_cast<uintptr_t>(a->workspace) & 255)
    return fail(TDM_EINVAL, "workspace must be 256-byte aligned");
  carve_d3(a->workspace, p->n, p->nbonds, p->ncells, p->cell_cap, p->kh, p->nblocks, w, &need);
  return TDM_OK;
}

// one evaluation at state (x, v) -> (x_out, v_out)
int launch_eval(tdm_d3_arrays *a, const tdm_d3_params *p, const D3Ws &w, int op, int flags,
                const double *x, const double *v, double *x_out, double *v_out, hipStream_t st)
{
  const int nb = p->nblocks;
  const int tab = int(a->evals & 1);
  const int hp = int(a->hist_flips & 1);
  const int upd_hist = p->friction && (flags & TDM_D3_UPDATE_HISTORY) && op != OP_ENERGY;
  StepArgs s;
  s.x = x;
  s.v = v;
  s.x_out = x_out;
  s.v_out = v_out;
  const size_t N = size_t(p->n), K = size_t(p->kh);
  s.hq_in = w.hq + size_t(hp) * N * K;
  s.hcnt_in = w.hcnt + size_t(hp) * N;
  s.hxi_in = w.hxi + size_t(hp) * N * K * 3;
  s.hq_out = w.hq + size_t(hp ^ 1) * N * K;
  s.hcnt_out = w.hcnt + size_t(hp ^ 1) * N;
  s.hxi_out = w.hxi + size_t(hp ^ 1) * N * K * 3;
  s.tab = tab;
  s.op = op;
  s.commit = (flags & TDM_D3_UPDATE_FRACTURE) && op != OP_ENERGY;
  s.upd_hist = upd_hist;
  hipLaunchKernelGGL(d3_strain_kernel, dim3(nb), dim3(kThreads), 0, st, *a, *p, w, x, tab);
  if (op != OP_ENERGY && p->nbonds > 0)
    hipLaunchKernelGGL(d3_bond_kernel, dim3((p->nbonds + kThreads - 1) / kThreads), dim3(kThreads),
                       0, st, *a, *p, w, x, s.commit);
  hipLaunchKernelGGL(d3_step_kernel, dim3(nb), dim3(kThreads), 0, st, *a, *p, w, s);
  if (op != OP_ENERGY)
    hipLaunchKernelGGL(d3_finish_kernel, dim3(1), dim3(kThreads), 0, st, *a, *p, w, op, nb);
  else
    hipLaunchKernelGGL(d3_status_kernel, dim3(1), dim3(64), 0, st, w, a->out);
  HIP_TRY(hipGetLastError());
  a->evals += 1;
  if (upd_hist) a->hist_flips += 1;
  return TDM_OK;
}

int launch_steps(tdm_d3_arrays *a, const tdm_d3_params *p, const D3Ws &w, int nsteps,
                 hipStream_t st)
{
  for (int k = 0; k < nsteps; ++k) {
    const bool even = (k & 1) == 0;
    int rc = launch_eval(a, p, w, OP_STEP, TDM_D3_UPDATE_FRACTURE | TDM_D3_UPDATE_HISTORY,
                         even ? a->x : a->x_alt, even ? a->v : a->v_alt,
                         even ? a->x_alt : a->x, even ? a->v_alt : a->v, st);
    if (rc) return rc;
    a->steps += 1;
  }
  if (nsteps & 1) {
    const size_t m = 3 * size_t(p->n);
    const int g = int((m + kThreads - 1) / kThreads) < 1024 ? int((m + kThreads - 1) / kThreads) : 1024;
    hipLaunchKernelGGL(d3_copy_kernel, dim3(g), dim3(kThreads), 0, st, a->x, a->x_alt, m);
    hipLaunchKernelGGL(d3_copy_kernel, dim3(g), dim3(kThreads), 0, st, a->v, a->v_alt, m);
    HIP_TRY(hipGetLastError());
  }
  return TDM_OK;
}

}  // namespace
}  // namespace tdm

using namespace tdm;

extern "C" {

int tdm_d3_abi_sizes(size_t *params_bytes, size_t *arrays_bytes)
{
  if (params_bytes) *params_bytes = sizeof(tdm_d3_params);
  if (arrays_bytes) *arrays_bytes = sizeof(tdm_d3_arrays);
  return TDM_OK;
}

int tdm_d3_blocks(int n) { return n > 0 ? (n + kThreads - 1) / kThreads : 0; }

size_t tdm_d3_workspace_bytes(int n, int nbonds, int ncells, int cell_cap, int kh)
{
  if (n < 1 || nbonds < 0 || ncells < 1 || cell_cap < 1 || kh < 1) return 0;
  D3Ws w;
  size_t total = 0;
  carve_d3(nullptr, n, nbonds, ncells, cell_cap, kh, tdm_d3_blocks(n), &w, &total);
  return total;
}

int tdm_d3_setup(tdm_d3_arrays *a, const tdm_d3_params *p, void *stream)
{
  D3Ws w;
  int rc = check_d3(a, p, &w);
  if (rc) return rc;
  hipStream_t st = static_cast<hipStream_t>(stream);
  const int nb = p->nblocks;
  hipLaunchKernelGGL(d3_covinv_kernel, dim3(nb), dim3(kThreads), 0, st, *a, *p, w);
  hipLaunchKernelGGL(fill_kernel<int>, dim3(256), dim3(kThreads), 0, st, w.count,
                     size_t(2) * p->ncells, 0);
  hipLaunchKernelGGL(fill_kernel<int>, dim3(64), dim3(kThreads), 0, st, w.hcnt,
                     size_t(2) * p->n, 0);
  hipLaunchKernelGGL(fill_kernel<int>, dim3(1), dim3(64), 0, st, w.status, size_t(64), 0);
  hipLaunchKernelGGL(fill_kernel<int>, dim3(1), dim3(64), 0, st, w.status + 2, size_t(1), 1);
  HIP_TRY(hipGetLastError());
  a->evals = 0;
  a->hist_flips = 0;
  a->steps = 0;
  return TDM_OK;
}

int tdm_d3_clear_status(tdm_d3_arrays *a, const tdm_d3_params *p, void *stream)
{
  D3Ws w;
  int rc = check_d3(a, p, &w);
  if (rc) return rc;
  hipLaunchKernelGGL(fill_kernel<int>, dim3(1), dim3(64), 0, static_cast<hipStream_t>(stream),
                     w.status, size_t(64), 0);
  HIP_TRY(hipGetLastError());
  return TDM_OK;
}

int tdm_d3_rebin(tdm_d3_arrays *a, const tdm_d3_params *p, void *stream)
{
  D3Ws w;
  int rc = check_d3(a, p, &w);
  if (rc) return rc;
  hipLaunchKernelGGL(fill_kernel<int>, dim3(1), dim3(64), 0, static_cast<hipStream_t>(stream),
                     w.status + 2, size_t(1), 1);
  HIP_TRY(hipGetLastError());
  return TDM_OK;
}

int tdm_d3_clear_history(tdm_d3_arrays *a, const tdm_d3_params *p, void *stream)
{
  D3Ws w;
  int rc = check_d3(a, p, &w);
  if (rc) return rc;
  hipLaunchKernelGGL(fill_kernel<int>, dim3(64), dim3(kThreads), 0,
                     static_cast<hipStream_t>(stream), w.hcnt, size_t(2) * p->n, 0);
  HIP_TRY(hipGetLastError());
  return TDM_OK;
}

int tdm_d3_step(tdm_d3_arrays *a, const tdm_d3_params *p, int nsteps, void *stream)
{
  D3Ws w;
  int rc = check_d3(a, p, &w);
  if (rc) return rc;
  if (nsteps < 0) return fail(TDM_EINVAL, "nsteps must be nonnegative");
  return launch_steps(a, p, w, nsteps, static_cast<hipStream_t>(stream));
}

int tdm_d3_forces(tdm_d3_arrays *a, const tdm_d3_params *p, int flags, void *stream)
{
  D3Ws w;
  int rc = check_d3(a, p, &w);
  if (rc) return rc;
  return launch_eval(a, p, w, OP_FORCES, flags, a->x, a->v, a->x_alt, a->v_alt,
                     static_cast<hipStream_t>(stream));
}

int tdm_d3_energy(tdm_d3_arrays *a, const tdm_d3_params *p, void *stream)
{
  D3Ws w;
  int rc = check_d3(a, p, &w);
  if (rc) return rc;
  if (!a->pair_energy) return fail(TDM_EINVAL, "pair_energy array missing");
  return launch_eval(a, p, w, OP_ENERGY, 0, a->x, a->v, a->x_alt, a->v_alt,
                     static_cast<hipStream_t>(stream));
}

int tdm_d3_shear(tdm_d3_arrays *a, const tdm_d3_params *p, double *shear, void *stream)
{
  D3Ws w;
  int rc = check_d3(a, p, &w);
  if (rc) return rc;
  if (!shear) return fail(TDM_EINVAL, "shear output missing");
  hipStream_t st = static_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(d3_strain_kernel, dim3(p->nblocks), dim3(kThreads), 0, st, *a, *p, w, a->x, -1);
  hipLaunchKernelGGL(d3_copy_kernel, dim3(64), dim3(kThreads), 0, st, shear, w.shear, size_t(p->n));
  HIP_TRY(hipGetLastError());
  return TDM_OK;
}

int tdm_d3_graph_create(tdm_d3_arrays *a, const tdm_d3_params *p, int pairs, void *stream,
                        tdm_graph **out)
{
  D3Ws w;
  int rc = check_d3(a, p, &w);
  if (rc) return rc;
  if (!out || pairs <= 0) return fail(TDM_EINVAL, "bad graph request");
  hipStream_t s = static_cast<hipStream_t>(stream);
  tdm_d3_arrays copy = *a;
  HIP_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  rc = launch_steps(&copy, p, w, 2 * pairs, s);
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

int tdm_d3_step_timed(tdm_d3_arrays *a, const tdm_d3_params *p, int nsteps, void *stream,
                      double *mean_step_us, double *mean_strain_us, double *mean_bond_us)
{
  D3Ws w;
  int rc = check_d3(a, p, &w);
  if (rc) return rc;
  if (nsteps <= 0 || (nsteps & 1)) return fail(TDM_EINVAL, "nsteps must be even and positive");
  hipStream_t st = static_cast<hipStream_t>(stream);
  hipEvent_t e[4];
  for (int k = 0; k < 4; ++k) HIP_TRY(hipEventCreate(&e[k]));
  double tot_step = 0.0, tot_strain = 0.0, tot_bond = 0.0;
  for (int k = 0; k < nsteps; ++k) {
    const bool even = (k & 1) == 0;
    const double *x = even ? a->x : a->x_alt, *v = even ? a->v : a->v_alt;
    double *xo = even ? a->x_alt : a->x, *vo = even ? a->v_alt : a->v;
    const int tab = int(a->evals & 1), hp = int(a->hist_flips & 1);
    const int upd = p->friction;
    StepArgs s;
    s.x = x; s.v = v; s.x_out = xo; s.v_out = vo;
    const size_t N = size_t(p->n), K = size_t(p->kh);
    s.hq_in = w.hq + size_t(hp) * N * K;
    s.hcnt_in = w.hcnt + size_t(hp) * N;
    s.hxi_in = w.hxi + size_t(hp) * N * K * 3;
    s.hq_out = w.hq + size_t(hp ^ 1) * N * K;
    s.hcnt_out = w.hcnt + size_t(hp ^ 1) * N;
    s.hxi_out = w.hxi + size_t(hp ^ 1) * N * K * 3;
    s.tab = tab; s.op = OP_STEP; s.commit = 1; s.upd_hist = upd;
    HIP_TRY(hipEventRecord(e[0], st));
    hipLaunchKernelGGL(d3_strain_kernel, dim3(p->nblocks), dim3(kThreads), 0, st, *a, *p, w, x, tab);
    HIP_TRY(hipEventRecord(e[1], st));
    if (p->nbonds > 0)
      hipLaunchKernelGGL(d3_bond_kernel, dim3((p->nbonds + kThreads - 1) / kThreads),
                         dim3(kThreads), 0, st, *a, *p, w, x, 1);
    HIP_TRY(hipEventRecord(e[3], st));
    hipLaunchKernelGGL(d3_step_kernel, dim3(p->nblocks), dim3(kThreads), 0, st, *a, *p, w, s);
    HIP_TRY(hipEventRecord(e[2], st));
    hipLaunchKernelGGL(d3_finish_kernel, dim3(1), dim3(kThreads), 0, st, *a, *p, w, int(OP_STEP),
                       p->nblocks);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipEventSynchronize(e[2]));
    float ms0 = 0.f, ms1 = 0.f, ms2 = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms0, e[0], e[1]));
    HIP_TRY(hipEventElapsedTime(&ms2, e[1], e[3]));
    HIP_TRY(hipEventElapsedTime(&ms1, e[3], e[2]));
    tot_strain += ms0;
    tot_bond += ms2;
    tot_step += ms1;
    a->evals += 1;
    if (upd) a->hist_flips += 1;
    a->steps += 1;
  }
  HIP_TRY(hipStreamSynchronize(st));
  for (int k = 0; k < 4; ++k) (void)hipEventDestroy(e[k]);
  if (mean_step_us) *mean_step_us = 1e3 * tot_step / nsteps;
  if (mean_strain_us) *mean_strain_us = 1e3 * tot_strain / nsteps;
  if (mean_bond_us) *mean_bond_us = 1e3 * tot_bond / nsteps;
  return TDM_OK;
}

}  // extern "C"
