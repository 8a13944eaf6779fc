// Connected components of a plate's surviving bonds on MI355X (gfx950):
// IceDEM.fragments (src/tensordem/dem.py:462-484).
//
// The reference walks the alive-bond graph breadth first from the lowest
// unvisited disk; each fragment lists its disks in ascending order, and the
// fragments come in the order of their lowest disk.  Here:
//   * a lock-free union-find over the lattice link bits hooks the larger root
//     under the smaller one, so every component ends with its lowest disk as
//     root, whatever order the hooks ran in;
//   * roots are numbered in ascending order by one block-wide scan, which also
//     gives each component's start in the output order;
//   * a stable LSD radix sort (8-bit digits, ranks inside a block from wave
//     ballots) groups the disk indices by component, ascending inside each;
//   * one wave per component sums positions and velocities in a fixed order
//     (lane-strided partial sums, then a fixed DPP tree), so the centroids and
//     mean velocities are deterministic; no float atomics anywhere.
#include "common.h"

#include <mutex>

namespace tdm {
namespace {

constexpr int kScanThreads = 1024;
constexpr int kSumBlocks = 2048;
constexpr int kMaxFragDisks = 1 << 24;  // three 8-bit radix passes

struct FragWs {
  int *parent;   // [n] union-find forest
  int *root;     // [n] root of each disk, then sort keys (ping)
  int *cnt;      // [n] disks per root
  int *rootcid;  // [n] component index of each root
  int *val;      // [n] sort values (ping)
  int *key2;     // [n] sort keys (pong)
  int *val2;     // [n] sort values (pong)
  int *hist;     // [256][nblk] digit counts, then their exclusive scan
};

__host__ __device__ inline int frag_blocks(int n) { return (n + kThreads - 1) / kThreads; }

__host__ __device__ inline size_t frag_carve(char *base, int n, FragWs *w)
{
  size_t off = 0;
  auto take = [&](size_t bytes) {
    char *p = base ? base + off : nullptr;
    off += align256(bytes);
    return reinterpret_cast<int *>(p);
  };
  const size_t nn = sizeof(int) * (size_t)n;
  w->parent = take(nn);
  w->root = take(nn);
  w->cnt = take(nn);
  w->rootcid = take(nn);
  w->val = take(nn);
  w->key2 = take(nn);
  w->val2 = take(nn);
  w->hist = take(sizeof(int) * 256 * (size_t)frag_blocks(n));
  return off;
}

__device__ inline int ld_relaxed(const int *p)
{
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Root of x with path halving.  Every parent pointer points to a smaller or
// equal index of the same set, so concurrent halving only ever shortens paths
// and the walk ends.
__device__ int find_root(int *parent, int x)
{
  int p = ld_relaxed(parent + x);
  while (p != x) {
    const int g = ld_relaxed(parent + p);
    if (g != p) __hip_atomic_store(parent + x, g, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    x = g;
    p = ld_relaxed(parent + x);
  }
  return x;
}

// Hook the larger root under the smaller; a failed CAS means that root was
// hooked meanwhile (its parent became smaller), so the retry starts lower.
__device__ void unite(int *parent, int a, int b)
{
  int ra = find_root(parent, a), rb = find_root(parent, b);
  while (ra != rb) {
    if (ra > rb) { const int t = ra; ra = rb; rb = t; }
    const int old = atomicCAS(parent + rb, rb, ra);
    if (old == rb) return;
    rb = find_root(parent, old);
    ra = find_root(parent, ra);
  }
}

__global__ void frag_init_kernel(FragWs w, int n)
{
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    w.parent[i] = i;
    w.cnt[i] = 0;
  }
}

// link bit c of disk i: the bond towards direction code c is alive; codes 4..7
// point to the partners above: +1, +nx-1, +nx, +nx+1
__global__ void frag_hook_kernel(const uint8_t *links, int nx, int ny, FragWs w)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nx * ny) return;
  const uint32_t l = links[i];
  if (!(l & 0xF0u)) return;
  const int ix = i % nx, iy = i / nx;
  if ((l >> 4 & 1u) && ix + 1 < nx) unite(w.parent, i, i + 1);
  if ((l >> 5 & 1u) && ix > 0 && iy + 1 < ny) unite(w.parent, i, i + nx - 1);
  if ((l >> 6 & 1u) && iy + 1 < ny) unite(w.parent, i, i + nx);
  if ((l >> 7 & 1u) && ix + 1 < nx && iy + 1 < ny) unite(w.parent, i, i + nx + 1);
}

__global__ void frag_root_kernel(FragWs w, int n)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int r = find_root(w.parent, i);
  w.root[i] = r;
  atomicAdd(w.cnt + r, 1);
}

// Exclusive block scan of one int per thread (kScanThreads threads); returns
// the thread's prefix and writes the total to *total (all threads).
__device__ int block_exclusive(int v, int *lds, int *total)
{
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  int incl = v;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const int o = __shfl_up(incl, d, 64);
    if (lane >= d) incl += o;
  }
  if (lane == 63) lds[wave] = incl;
  __syncthreads();
  if (t == 0) {
    int run = 0;
    for (int k = 0; k < kScanThreads / 64; ++k) {
      const int s = lds[k];
      lds[k] = run;
      run += s;
    }
    lds[kScanThreads / 64] = run;
  }
  __syncthreads();
  const int out = lds[wave] + incl - v;
  *total = lds[kScanThreads / 64];
  __syncthreads();
  return out;
}

// Roots in ascending order get component numbers 0, 1, ...; component c
// starts in the output order at the number of disks of the roots below it.
__global__ __launch_bounds__(kScanThreads) void frag_number_kernel(FragWs w, int n,
                                                                   int *comp_first, int *ncomp)
{
  __shared__ int lds[kScanThreads / 64 + 1];
  const int per = (n + kScanThreads - 1) / kScanThreads;
  const int lo = min(n, (int)threadIdx.x * per), hi = min(n, lo + per);
  int roots = 0, disks = 0;
  for (int i = lo; i < hi; ++i) {
    if (w.root[i] == i) ++roots;
    disks += w.cnt[i];
  }
  int nroots, ndisks;
  int r0 = block_exclusive(roots, lds, &nroots);
  int d0 = block_exclusive(disks, lds, &ndisks);
  for (int i = lo; i < hi; ++i) {
    if (w.root[i] == i) {
      w.rootcid[i] = r0;
      comp_first[r0] = d0;
      ++r0;
    }
    d0 += w.cnt[i];
  }
  if (threadIdx.x == 0) {
    *ncomp = nroots;
    comp_first[nroots] = ndisks;
  }
}

__global__ void frag_label_kernel(FragWs w, int n, int *labels)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int c = w.rootcid[w.root[i]];
  labels[i] = c;
  w.root[i] = c;  // sort key
  w.val[i] = i;   // sort value
}

// digit counts of one block's 256 keys: hist[d][block]
__global__ __launch_bounds__(kThreads) void frag_hist_kernel(const int *key, int n, int shift,
                                                             int *hist)
{
  __shared__ int h[256];
  const int t = threadIdx.x, i = blockIdx.x * kThreads + t;
  h[t] = 0;
  __syncthreads();
  if (i < n) atomicAdd(h + ((key[i] >> shift) & 255), 1);
  __syncthreads();
  hist[(size_t)t * gridDim.x + blockIdx.x] = h[t];
}

// in-place exclusive scan of m ints by one block (digit-major histogram)
__global__ __launch_bounds__(kScanThreads) void frag_scan_kernel(int *a, int m)
{
  __shared__ int lds[kScanThreads / 64 + 1];
  const int per = (m + kScanThreads - 1) / kScanThreads;
  const int lo = min(m, (int)threadIdx.x * per), hi = min(m, lo + per);
  int s = 0;
  for (int i = lo; i < hi; ++i) s += a[i];
  int total;
  int run = block_exclusive(s, lds, &total);
  for (int i = lo; i < hi; ++i) {
    const int v = a[i];
    a[i] = run;
    run += v;
  }
}

// Stable scatter of one block's 256 (key, value) pairs: the rank of an
// element among the block's elements with the same digit is its rank inside
// its wave (ballot match on the 8 digit bits) plus the counts of the waves
// before it.
__global__ __launch_bounds__(kThreads) void frag_scatter_kernel(const int *key, const int *val,
                                                                int n, int shift,
                                                                const int *hist, int *key_out,
                                                                int *val_out)
{
  __shared__ int cnt[kThreads / 64][256];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int i = blockIdx.x * kThreads + t;
  for (int k = 0; k < kThreads / 64; ++k) cnt[k][t] = 0;
  const bool valid = i < n;
  const int k_in = valid ? key[i] : 0;
  const int d = (k_in >> shift) & 255;
  unsigned long long peers = __ballot(valid);
#pragma unroll
  for (int b = 0; b < 8; ++b) {
    const unsigned long long m = __ballot(valid && ((d >> b) & 1));
    peers &= ((d >> b) & 1) ? m : ~m;
  }
  const unsigned long long below = (1ull << lane) - 1ull;
  const int rank = __popcll(peers & below);
  __syncthreads();
  if (valid && rank == 0) cnt[wave][d] = __popcll(peers);
  __syncthreads();
  if (valid) {
    int pre = 0;
    for (int k = 0; k < wave; ++k) pre += cnt[k][d];
    const int dst = hist[(size_t)d * gridDim.x + blockIdx.x] + pre + rank;
    key_out[dst] = k_in;
    val_out[dst] = val[i];
  }
}

// One wave per component: sums of x, y, vx, vy over its disks in ascending
// order, lane-strided, then a fixed DPP tree.
__global__ __launch_bounds__(kThreads) void frag_sum_kernel(const double2 *x, const double2 *v,
                                                            const int *order,
                                                            const int *comp_first,
                                                            const int *ncomp, double *sums)
{
  const int lane = threadIdx.x & 63;
  const int waves = gridDim.x * (kThreads / 64);
  const int nc = *ncomp;
  for (int c = blockIdx.x * (kThreads / 64) + (threadIdx.x >> 6); c < nc; c += waves) {
    const int s = comp_first[c], e = comp_first[c + 1];
    double sx = 0.0, sy = 0.0, svx = 0.0, svy = 0.0;
    for (int k = s + lane; k < e; k += 64) {
      const int p = order[k];
      const double2 xp = x[p], vp = v[p];
      sx += xp.x;
      sy += xp.y;
      svx += vp.x;
      svy += vp.y;
    }
    sx = wave_sum(sx);
    sy = wave_sum(sy);
    svx = wave_sum(svx);
    svy = wave_sum(svy);
    if (lane == 0) {
      double *o = sums + 4 * (size_t)c;
      o[0] = sx;
      o[1] = sy;
      o[2] = svx;
      o[3] = svy;
    }
  }
}

int check_fragments(const uint8_t *links, const double *x, const double *v, int nx, int ny,
                    int32_t *labels, int32_t *order, int32_t *comp_first, int32_t *ncomp,
                    double *comp_sums)
{
  if (nx < 1 || ny < 1 || (long long)nx * ny > kMaxFragDisks)
    return fail(TDM_EINVAL, "fragments: lattice must hold 1 to 2^24 disks");
  if (!links || !x || !v || !labels || !order || !comp_first || !ncomp || !comp_sums)
    return fail(TDM_EINVAL, "fragments: null array");
  return TDM_OK;
}

// The kernel sequence of both fragment entry points; `ws` holds
// frag_carve(n) bytes, 256-byte aligned.
int launch_fragments(const uint8_t *links, const double *x, const double *v, int nx, int ny,
                     int32_t *labels, int32_t *order, int32_t *comp_first, int32_t *ncomp,
                     double *comp_sums, char *ws, hipStream_t s)
{
  const int n = nx * ny;
  FragWs w;
  frag_carve(ws, n, &w);
  const int nb = frag_blocks(n);
  frag_init_kernel<<<min(nb, 1024), kThreads, 0, s>>>(w, n);
  frag_hook_kernel<<<nb, kThreads, 0, s>>>(links, nx, ny, w);
  frag_root_kernel<<<nb, kThreads, 0, s>>>(w, n);
  frag_number_kernel<<<1, kScanThreads, 0, s>>>(w, n, comp_first, ncomp);
  frag_label_kernel<<<nb, kThreads, 0, s>>>(w, n, labels);
  // component numbers are below n: one 8-bit pass per byte of n - 1
  int passes = 1;
  while (passes < 3 && ((n - 1) >> (8 * passes)) != 0) ++passes;
  int *kin = w.root, *vin = w.val;
  for (int p = 0; p < passes; ++p) {
    const bool last = p == passes - 1;
    int *kout = (kin == w.root) ? w.key2 : w.root;
    int *vout = last ? order : ((vin == w.val) ? w.val2 : w.val);
    frag_hist_kernel<<<nb, kThreads, 0, s>>>(kin, n, 8 * p, w.hist);
    frag_scan_kernel<<<1, kScanThreads, 0, s>>>(w.hist, 256 * nb);
    frag_scatter_kernel<<<nb, kThreads, 0, s>>>(kin, vin, n, 8 * p, w.hist, kout, vout);
    kin = kout;
    vin = vout;
  }
  frag_sum_kernel<<<min(kSumBlocks, nb), kThreads, 0, s>>>(
      reinterpret_cast<const double2 *>(x), reinterpret_cast<const double2 *>(v), order,
      comp_first, ncomp, comp_sums);
  HIP_TRY(hipGetLastError());
  return TDM_OK;
}

// ---------------------------------------------------------------------------
// Allocation-free variant: the scratch lives in a file-scope device array of
// a fixed size (each GPU the library runs on holds its own copy until the
// process ends).  In host code `scratch` names a host-side stand-in, not the
// device address: the address comes from hipGetSymbolAddress, once per
// device, with that device current, and only that pointer is passed on.
// ---------------------------------------------------------------------------
constexpr int kNoAllocMaxDisks = 1 << 18;  // 262 144 disks (a 512 x 512 plate)
constexpr size_t kIntArrayBytes = ((sizeof(int) * (size_t)kNoAllocMaxDisks + 255) / 256) * 256;
constexpr size_t kHistBytes =
    ((sizeof(int) * 256 * (size_t)((kNoAllocMaxDisks + kThreads - 1) / kThreads) + 255) / 256) *
    256;
constexpr size_t kNoAllocBytes = 7 * kIntArrayBytes + kHistBytes;  // 8 MiB
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

std::mutex g_scratch_mutex;
void *g_scratch_ptr[kMaxDevices] = {};

// Device address of this device's scratch array (looked up on the first call
// on each device and kept).
int scratch_address(char **out)
{
  int dev = -1;
  HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(TDM_EINVAL, "fragments: device id out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[dev]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(TDM_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[dev] = p;
  }
  *out = static_cast<char *>(g_scratch_ptr[dev]);
  return TDM_OK;
}

}  // namespace
}  // namespace tdm

using namespace tdm;

extern "C" {

size_t tdm_fragments_workspace_bytes(int n)
{
  if (n < 1) return 0;
  FragWs w;
  return frag_carve(nullptr, n, &w);
}

int tdm_plate_fragments(const uint8_t *links, const double *x, const double *v, int nx, int ny,
                        int32_t *labels, int32_t *order, int32_t *comp_first, int32_t *ncomp,
                        double *comp_sums, void *workspace, size_t workspace_bytes,
                        void *stream)
{
  int rc = check_fragments(links, x, v, nx, ny, labels, order, comp_first, ncomp, comp_sums);
  if (rc) return rc;
  if (!workspace || (reinterpret_cast<uintptr_t>(workspace) & 255))
    return fail(TDM_EINVAL, "fragments: workspace must be non-null and 256-byte aligned");
  if (workspace_bytes < tdm_fragments_workspace_bytes(nx * ny))
    return fail(TDM_ESPACE, "fragments: workspace too small");
  return launch_fragments(links, x, v, nx, ny, labels, order, comp_first, ncomp, comp_sums,
                          static_cast<char *>(workspace), static_cast<hipStream_t>(stream));
}

int tdm_plate_fragments_noalloc(const uint8_t *links, const double *x, const double *v, int nx,
                                int ny, int32_t *labels, int32_t *order, int32_t *comp_first,
                                int32_t *ncomp, double *comp_sums, void *stream)
{
  int rc = check_fragments(links, x, v, nx, ny, labels, order, comp_first, ncomp, comp_sums);
  if (rc) return rc;
  if ((long long)nx * ny > kNoAllocMaxDisks)
    return fail(TDM_EINVAL, "fragments_noalloc: at most 262144 disks (nx * ny) per call");
  if (tdm_fragments_workspace_bytes(nx * ny) > kNoAllocBytes)
    return fail(TDM_ESPACE, "fragments_noalloc: scratch smaller than the workspace layout");
  char *ws = nullptr;
  rc = scratch_address(&ws);
  if (rc) return rc;
  return launch_fragments(links, x, v, nx, ny, labels, order, comp_first, ncomp, comp_sums, ws,
                          static_cast<hipStream_t>(stream));
}

}  // extern "C"
