/*
 * tensordem_amd: MI355X (gfx950) kernels for the force-and-integrate step of
 * the 2-D bonded-disk ice solver, behind a plain C ABI.
 *
 * The reference solver is pure Python/PyTorch and exposes no foreign-function
 * interface; the entry points below replace the bodies of its Python methods
 * and are bound from Python with ctypes (tensordem_amd/_lib.py).  Every
 * pointer argument is a device pointer (HIP global memory) unless the comment
 * says "host".  Arrays are owned by the caller; the library allocates nothing
 * except graph objects inside a tdm_graph.  `stream` is a hipStream_t
 * (NULL = default stream).  Every function returns TDM_OK (0) or a negative
 * TDM_E* code; tdm_last_error() describes the most recent failure.
 *
 * Particle layout: one lattice ("member") of nx*ny disks is stored row-major
 * (particle = first + iy*nx + ix), exactly the order of the reference
 * (src/tensordem/dem.py:105-110).  Bonds of a member are its 8-neighbour
 * lattice links in the reference's triu order (dem.py:116-124); a bond is
 * identified by (lower particle, direction code 4..7) and its global index is
 * fwd_bond[lower] + rank of the code among the lower particle's links.
 */
#ifndef TENSORDEM_AMD_H
#define TENSORDEM_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TDM_OK 0
#define TDM_EINVAL (-1)   /* bad argument or shape */
#define TDM_EHIP (-2)     /* a HIP runtime call failed */
#define TDM_ESPACE (-3)   /* workspace too small */

/* status bits accumulated per member by the kernels (tdm_plate_member.status) */
#define TDM_ST_NONFINITE 1      /* x or v became non-finite (dem.py:379-382) */
#define TDM_ST_STALE_LIST 2     /* a disk moved >= skin/2 since the contact list was built */
#define TDM_ST_CAND_OVERFLOW 4  /* more unbonded contact candidates than slots */
#define TDM_ST_CELL_OVERFLOW 8  /* more disks in one hash cell than its capacity */
#define TDM_ST_HIST_OVERFLOW 16 /* more friction springs on one sphere than history slots */
#define TDM_ST_INTEGRATOR 32    /* a velocity-Verlet bed member stepped by a batch whose
                                   tdm_bed_arrays.verlet is 0 (it was stepped by Euler) */

/* ledger layout (accounting, per member, float64) -- the dict of dem.py:139-144 */
#define TDM_ACC_EXT_WORK 0
#define TDM_ACC_RELEASE 1
#define TDM_ACC_ACQUIRED 2
#define TDM_ACC_TOOL_WORK 3
#define TDM_ACC_DAMPING 4
#define TDM_ACC_EXT_IMP 5   /* x, y */
#define TDM_ACC_SUP_IMP 7   /* x, y */
#define TDM_ACC_DRAG_IMP 9  /* x, y */
#define TDM_ACC_TOOL_IMP 11 /* x, y */
#define TDM_NACC 13

/* last-step record (per member, float64) -- the dict of dem.py:372-376 */
#define TDM_LAST_T0 0
#define TDM_LAST_T1 1
#define TDM_LAST_WORK 2
#define TDM_LAST_SUPPORT 3  /* x, y */
#define TDM_LAST_REACTION 5 /* x, y */
#define TDM_LAST_EXTERNAL 7 /* x, y */
#define TDM_LAST_MOMENT 9
#define TDM_NLAST 10

/*
 * One plate.  Lives in device memory (array of nmembers); the step kernel
 * advances `time`, `step_count`, `mode` and `tab` and ORs into `status`.
 * Field meaning follows DEMConfig (dem.py:17-35) unless noted.
 */
typedef struct tdm_plate_member {
  double radius, bond_stiffness, contact_stiffness;
  double breaking_strain, shear_breaking_strain;
  double contact_damping, drag, tool_radius;
  double dt, mass;
  double tool_start[2];   /* circle: tool start; hull: hull_start */
  double tool_velocity[2];/* circle: (0,-tool_speed); hull: hull_velocity */
  double time;            /* simulated time at the start of the next step */
  double skin;            /* Verlet skin for unbonded contact candidates [m] */
  double cell_size;       /* hash cell edge, >= 2*radius + skin [m] */
  double grid_origin[2];  /* lower-left corner of the member's hash grid [m] */
  double fp32_margin;     /* relative band around a break threshold re-checked in FP64 */
  int64_t step_count;
  int32_t nx, ny;         /* lattice size */
  int32_t first;          /* first particle index */
  int32_t bond_first;     /* first global bond index */
  int32_t nbonds;
  int32_t hull_first, hull_m; /* hull vertices [hull_first, +hull_m) in `hull`; 0 = circle */
  int32_t ncx, ncy, cell_first; /* hash grid size and first cell index */
  int32_t tile_x;         /* tile width in disks (16..256); tile height = 256 / tile_x */
  int32_t block_first, nblocks; /* this member's range in the block table */
  int32_t mode;           /* 0 idle, 1 search candidates this step, 2 bin new positions */
  int32_t tab;            /* hash table holding the current candidate source */
  int32_t status;         /* TDM_ST_* bits */
  int32_t has_loads;      /* loads array rows for this member are valid */
  int32_t per_disk_strength; /* eps_t/eps_s come from the per-disk arrays */
  int32_t pad_;
} tdm_plate_member;

/* Device arrays of a plate batch (all members share them). */
typedef struct tdm_plate_arrays {
  double *x, *v;            /* (N,2) state, updated by tdm_plate_step */
  double *x_alt, *v_alt;    /* (N,2) second buffers (double buffering) */
  const uint8_t *fixed;     /* (N,) clamped disks (dem.py:113-115) */
  uint8_t *links;           /* (N,) alive bit per direction code 0..7 */
  const int32_t *fwd_bond;  /* (N,) first global bond index owned by the disk */
  uint8_t *alive;           /* (B,) dem.py:130 (bond-indexed) */
  uint8_t *failure_mode;    /* (B,) 0 intact, 1 tensile, 2 shear, 3 both (dem.py:131-133) */
  double *failure_time_s, *failure_extension_m, *failure_energy_J; /* (B,) */
  const double *loads;      /* (N,2) external forces or NULL (dem.py:210-224) */
  const double *hull;       /* (H,2) hull profile vertices or NULL (dem.py:154-169) */
  const double *eps_t_disk, *eps_s_disk; /* (N,) per-disk thresholds or NULL */
  double *accounting;       /* (nmembers, TDM_NACC) */
  double *last;             /* (nmembers, TDM_NLAST) */
  double *force;            /* (N,2) output of tdm_plate_forces, else may be NULL */
  double *reaction;         /* (nmembers,2) output of tdm_plate_forces */
  double *energy;           /* (nmembers,5) output of tdm_plate_energy */
  tdm_plate_member *members;/* (nmembers,) */
  void *workspace;          /* tdm_plate_workspace_bytes() bytes, 256-byte aligned */
  size_t workspace_bytes;
  int32_t n, nbonds, nmembers, nblocks, ncells;
  int32_t cand_slots;       /* unbonded candidate slots per disk */
  int32_t cell_cap;         /* disks per hash cell */
  int32_t max_tile_x;       /* widest member tile, or 0 if not given; a batch whose
                               tiles are at most 128 wide may stage candidate forces
                               in LDS (checked against the members at setup) */
} tdm_plate_arrays;

/* Library version (major*10000 + minor*100 + patch). */
int tdm_version(void);

/* Copies the last error message into buf (host), NUL-terminated. */
int tdm_last_error(char *buf, size_t len);

/* sizeof(tdm_plate_member) and sizeof(tdm_plate_arrays) as compiled, so a
 * binding can check its mirror of the two structs. */
int tdm_abi_sizes(size_t *member_bytes, size_t *arrays_bytes);

/* Number of 256-disk tiles (thread blocks) a member needs for tile width tile_x. */
int tdm_plate_tiles(int nx, int ny, int tile_x);

/* Workspace bytes for a batch (block table, ledger partials, counters, hash
 * tables, candidate lists). */
size_t tdm_plate_workspace_bytes(int n, int nblocks, int nmembers, int ncells, int cand_slots,
                                 int cell_cap);

/*
 * Prepares a batch: writes the XCD-aware block table into the workspace,
 * derives per-disk link bits from `alive`, clears counters and hash tables
 * and bins the current positions so the next step builds fresh contact
 * candidate lists.  `members` must already hold the member table (device).
 * Replaces the topology part of IceDEM.__init__ (dem.py:105-153).
 */
int tdm_plate_setup(const tdm_plate_arrays *a, void *stream);

/* Re-derives link bits from `alive` after the caller edited `alive`
 * (the reference lets callers assign IceDEM.alive directly). */
int tdm_plate_sync_links(const tdm_plate_arrays *a, void *stream);

/* Re-bins the current positions and schedules a candidate search; call after
 * the caller moved disks (assigned positions). */
int tdm_plate_rebin(const tdm_plate_arrays *a, void *stream);

/*
 * Advances every member by `nsteps` steps of IceDEM.step (dem.py:347-382):
 * fracture commit, bond/contact/tool/hull forces, semi-implicit Euler, clamp,
 * impulse/work/energy ledger.  State alternates between (x,v) and
 * (x_alt,v_alt); after an odd nsteps the new state is in the _alt buffers.
 * fp32 != 0 evaluates bond, contact and tool forces in FP32 (state, sums and
 * ledger stay FP64; break decisions within fp32_margin of a threshold are
 * re-evaluated in FP64).  `parity` (0/1) says which buffer holds the state on
 * entry (0 = x/v).
 */
int tdm_plate_step(const tdm_plate_arrays *a, int nsteps, int parity, int fp32, void *stream);

/*
 * tdm_plate_step for nsteps steps (parity 0 on entry) with a HIP event pair
 * attached to every step-kernel dispatch on `stream` (hipExtLaunchKernelGGL;
 * the ledger launch is not timed); writes the mean step-kernel duration in
 * microseconds.  Measurement helper with no counterpart in the reference.
 */
int tdm_plate_step_timed(const tdm_plate_arrays *a, int nsteps, int fp32, void *stream,
                         double *mean_kernel_us);

/*
 * IceDEM.forces (dem.py:226-345): commits fracture, writes `force` (N,2) and
 * the ice-on-tool `reaction` (nmembers,2); positions and time do not move.
 */
int tdm_plate_forces(const tdm_plate_arrays *a, int parity, int fp32, void *stream);

/*
 * IceDEM.mechanical_energy (dem.py:603-624): writes per member
 * (kinetic, bond, pair contact, tool contact, total) into `energy`; commits
 * nothing.
 */
int tdm_plate_energy(const tdm_plate_arrays *a, int parity, void *stream);

/*
 * IceDEM.fragments (dem.py:462-484): connected components of one plate's
 * surviving bonds, isolated disks included.  `links` (nx*ny,) are the plate's
 * link bits (tdm_plate_arrays.links from the plate's first disk: bit c set
 * when the bond towards direction code c is alive, codes 4..7 pointing to
 * +1, +nx-1, +nx, +nx+1), `x` / `v` its (nx*ny,2) positions and velocities.
 * Outputs (device, int32 unless noted):
 *   labels (n,): component of each disk; components are numbered in the
 *     order of their lowest disk, the order the reference lists them in;
 *   order (n,): disk indices grouped by component, ascending inside each
 *     (the reference's sorted particle_ids, concatenated);
 *   comp_first (n+1,): component c is order[comp_first[c] .. comp_first[c+1]);
 *   ncomp (1,): number of components;
 *   comp_sums (n,4) float64: per component the sums of x, y, vx, vy over its
 *     disks (centroid and mean velocity are these over the disk count).
 * Deterministic: sums in a fixed order, no float atomics.  `workspace` holds
 * tdm_fragments_workspace_bytes(nx*ny) bytes, 256-byte aligned; n <= 2^24.
 */
size_t tdm_fragments_workspace_bytes(int n);
int tdm_plate_fragments(const uint8_t *links, const double *x, const double *v, int nx, int ny,
                        int32_t *labels, int32_t *order, int32_t *comp_first, int32_t *ncomp,
                        double *comp_sums, void *workspace, size_t workspace_bytes,
                        void *stream);

/*
 * tdm_plate_fragments without a workspace argument and without allocating:
 * its scratch is a file-scope device array of 8 MiB (7 int arrays and the
 * radix histogram for up to 262 144 disks), which each GPU the library runs
 * on holds until the process ends.  No counterpart in the reference; it is
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * outputs, bit for bit, as tdm_plate_fragments; nx * ny above 262 144 is
 * rejected with TDM_EINVAL.  Every call on one GPU shares that GPU's array:
 * calls that overlap on one GPU (concurrent streams or host threads) race.
 * `stream` must belong to the current device.
 */
int tdm_plate_fragments_noalloc(const uint8_t *links, const double *x, const double *v, int nx,
                                int ny, int32_t *labels, int32_t *order, int32_t *comp_first,
                                int32_t *ncomp, double *comp_sums, void *stream);

/* Opaque captured sequence of steps (a hipGraph of 2*pairs steps). */
typedef struct tdm_graph tdm_graph;

/* Captures 2*pairs steps starting from parity 0 into a graph (the state ends
 * in x/v again).  Kernel arguments are frozen at capture time. */
int tdm_plate_graph_create(const tdm_plate_arrays *a, int pairs, int fp32, void *stream,
                           tdm_graph **out);
int tdm_graph_launch(tdm_graph *g, void *stream);
int tdm_graph_destroy(tdm_graph *g);

/* ------------------------------------------------------------------------
 * Unbonded Hertz-Mindlin disk beds (2-D restatement of the reference's
 * floe-bed law, src/tensordem/scale_ceiling/kernel.py:277-401): Hertz normal
 * force with a restitution-derived viscous term, Cundall-Strack tangential
 * spring capped by Coulomb friction with per-pair history that is dropped on
 * separation, floor and back-wall contacts, gravity, semi-implicit Euler.
 * Disks of a member are contiguous; any order works, spatially coherent
 * order is fastest.
 * ---------------------------------------------------------------------- */

/* per-member bed ledger (accounting), float64 */
#define TDM_BED_DASHPOT 0      /* energy absorbed by the pair viscous terms [J] */
#define TDM_BED_FRICTION 1     /* work of the pair tangential springs [J] */
#define TDM_BED_WALL_DASHPOT 2 /* ... wall viscous terms [J] */
#define TDM_BED_WALL_FRICTION 3
#define TDM_BED_WALL_IMPULSE 4 /* x, y: impulse of the walls on the disks [N s] */
#define TDM_BED_CONTACTS 6     /* pair contacts during the last step */
#define TDM_BED_WALL_CONTACTS 7
#define TDM_BED_WALL_WORK 8    /* work of the moving right wall on the disks [J] */
#define TDM_BED_NACC 9

typedef struct tdm_bed_member {
  double e_star, g_star;       /* pair moduli E/(2(1-nu^2)), G/(2(2-nu)) */
  double damping;              /* 2 sqrt(5/6) beta */
  double friction, wall_friction;
  double rho_pi_t;             /* mass = rho_pi_t * r^2 */
  double gravity_y;            /* [m/s^2], e.g. -9.81 */
  double dt;
  double floor_y, back_wall_x;
  /* right wall (gate or indenter, scale_ceiling/kernel.py:156-197, 338-401):
   * at x = right_wall_x + right_wall_v * time during the step that starts at
   * `time`, inward normal -x, moving with (right_wall_v, 0) */
  double right_wall_x, right_wall_v;
  double time, skin, cell_size, grid_origin[2];
  int64_t step_count;
  int32_t first, n;            /* disks [first, first+n) */
  int32_t use_floor, use_back_wall;
  int32_t ncx, ncy, cell_first;
  int32_t block_first, nblocks;
  int32_t mode, tab, list_parity, status;
  int32_t rebuilds;             /* candidate-list rebuilds so far */
  int32_t use_right_wall;
  /* integrator: 0 semi-implicit Euler (v and x are the state after the last
   * step); 1 the floe bed's kick-drift-kick velocity-Verlet
   * (scale_ceiling/kernel.py:410-421), one force evaluation per step: a step
   * evaluates the force at the current x with the current v, kicks v by
   * dt/2 F/m twice (the second half of the previous step, then the first
   * half of this one) and drifts x by dt v.  v is then the half-step velocity
   * the law sees; tdm_bed_velocity gives the full-step one.  verlet_start = 1
   * marks v as a full-step velocity (a new or edited state: the first step
   * skips the first kick, as the reference's first step evaluates the force
   * of the state it is given); the steps clear it. */
  int32_t integrator, verlet_start, pad_;
} tdm_bed_member;

typedef struct tdm_bed_arrays {
  double *x, *v, *x_alt, *v_alt; /* (N,2) state and second buffers */
  const double *radius;          /* (N,) */
  double *xi_wall;               /* (N,6) floor, back-wall and right-wall spring history */
  double *accounting;            /* (nmembers, TDM_BED_NACC) */
  double *energy;                /* (nmembers,4): kinetic, gravity, Hertz normal, tangential */
  tdm_bed_member *members;
  void *workspace;
  size_t workspace_bytes;
  int32_t n, nmembers, nblocks, ncells, cand_slots, cell_cap;
  int32_t fp32_histories;        /* 1: the spring histories in the workspace are float2 (a bed
                                    stepped with FP32 forces), 0: double2; every step call
                                    must pass the matching fp32 flag */
  int32_t verlet;                /* nonzero iff some member has integrator 1: such batches
                                    run kernels that carry the velocity-Verlet update (the
                                    others do not pay for it); setup rejects a Verlet member
                                    in a batch without the flag */
} tdm_bed_arrays;

/* sizeof(tdm_bed_member) and sizeof(tdm_bed_arrays) as compiled. */
int tdm_bed_abi_sizes(size_t *member_bytes, size_t *arrays_bytes);
/* Blocks (256 disks each) a member of n disks occupies. */
int tdm_bed_blocks(int n);
size_t tdm_bed_workspace_bytes(int n, int nblocks, int nmembers, int ncells, int cand_slots,
                               int cell_cap);
/* Byte offsets inside the workspace of the candidate lists (int32
 * [2][slots][N]), their spring histories (double2 [2][slots][N]), the
 * contact masks (uint32 [2][N]: bit e = slot e in contact, bits 16/17 =
 * floor/back wall) and counts (uint8 [2][N]), index 0/1 being the list
 * parity; and of each disk's position when its list was built (double2 [N]).
 * With fp32_histories the histories are float2 [2][slots][N] at the same
 * offset. */
int tdm_bed_workspace_layout(int n, int nblocks, int nmembers, int ncells, int cand_slots,
                             int cell_cap, size_t *cand_off, size_t *xi_off, size_t *mask_off,
                             size_t *count_off, size_t *xb_off);
/* Block table, cleared histories, first binning (members filled by caller).
 * Replaces FloeBed.__init__ state setup (scale_ceiling/kernel.py:158-216). */
int tdm_bed_setup(const tdm_bed_arrays *a, void *stream);
/* Re-bins x and schedules a candidate search (after the caller moved disks);
 * replaces FloeBed.build_neighbours (scale_ceiling/kernel.py:230-266). */
int tdm_bed_rebin(const tdm_bed_arrays *a, void *stream);
/* nsteps steps of the bed law; state alternates x/v <-> x_alt/v_alt starting
 * from x/v (parity 0).  fp32 != 0: forces in FP32, contact detection, state,
 * histories and ledger in FP64.  Replaces FloeBed.step / forces
 * (scale_ceiling/kernel.py:277-452); list rebuilds with history carry-over
 * happen inside (scale_ceiling/kernel.py:250-275). */
int tdm_bed_step(const tdm_bed_arrays *a, int nsteps, int fp32, void *stream);
/* As tdm_bed_step (nsteps even) with HIP events attached to each step-kernel
 * dispatch; mean step-kernel duration in microseconds.  Measurement helper. */
int tdm_bed_step_timed(const tdm_bed_arrays *a, int nsteps, int fp32, void *stream,
                       double *mean_kernel_us);
/* Kinetic, gravitational (zero at y = 0), Hertz normal and tangential spring
 * energy of the current state into `energy`; replaces FloeBed.kinetic_energy /
 * gravitational_energy / contact_energy (scale_ceiling/kernel.py:455-482). */
int tdm_bed_energy(const tdm_bed_arrays *a, void *stream);
/* Full-step velocity of the current state into v_alt (N,2): for
 * velocity-Verlet members v + (dt/2 F) / m with F the force of the current
 * x and v (the floe bed's `vel` after a step, scale_ceiling/kernel.py:
 * 417-418), springs and ledger untouched; for the others, and for a state
 * marked verlet_start, v itself.  The next step overwrites v_alt. */
int tdm_bed_velocity(const tdm_bed_arrays *a, int fp32, void *stream);
/* 2*pairs bed steps captured as a graph (state ends in x/v). */
int tdm_bed_graph_create(const tdm_bed_arrays *a, int pairs, int fp32, void *stream,
                         tdm_graph **out);

/* ------------------------------------------------------------------------
 * Rigid frictional disks with rotation (the reference's DiskSystem,
 * src/tensordem/granular2d/engine.py:144-684): linear spring-dashpot normal
 * force, tangential spring projected on the contact plane and capped by
 * Coulomb friction, rolling resistance, static wall and hull segments
 * (closest point on each segment), optional wall velocities, gravity, drag,
 * velocity decay, periodic axes, semi-implicit Euler; FP64 throughout, each
 * disk's sums in the reference's order, so trajectories match it bit for bit.
 * ---------------------------------------------------------------------- */

/* `out` (float64), written by every step or inspection */
#define TDM_DISK_NORMAL_DASHPOT 0 /* ledger, accumulated while enabled [J] */
#define TDM_DISK_FRICTION 1       /* ledger: minus the Coulomb sliding loss [J] */
#define TDM_DISK_ROLLING 2        /* ledger: minus the rolling plastic loss [J] */
#define TDM_DISK_ELASTIC 3        /* normal, tangential, rolling spring energy [J] */
#define TDM_DISK_WALL_FORCE 6     /* x, y: force of the disks on the walls [N] */
#define TDM_DISK_HULL_FORCE 8     /* x, y: force of the disks on the hull [N] */
#define TDM_DISK_CONTACTS 10      /* pair, wall, hull contacts of the evaluation */
#define TDM_DISK_STATUS 13        /* TDM_ST_* bits */
#define TDM_DISK_REBUILDS 14      /* candidate-list rebuilds so far */
#define TDM_DISK_CANDIDATES 15    /* candidate-list entries of the evaluation */
#define TDM_DISK_NOUT 16
/* per-contact record fields of tdm_disk_arrays.records */
#define TDM_DISK_NREC 14 /* i, j, kind, n(2), lever(2), f(2), fn, ft, overlap, point(2) */

typedef struct tdm_disk_params {
  double kn, zeta2, kt;        /* kn, 2*zeta, kt_ratio*kn */
  double mu, mu_rolling, kr_coef; /* kr_coef = kt/4 (k_r = kr_coef * r_eff^2) */
  double dt;
  double gravity[2], drift[2];
  double drag_time;            /* 0: no drag */
  double decay;                /* exp(-dt/damping_time), used when use_decay */
  double period[2];            /* > 0: periodic axis of that length */
  double reach;                /* candidate distance 2*max(r) + skin */
  double skin;                 /* rebuild when a disk moved more than skin/2 */
  double cell_h[2], grid_origin[2];
  int32_t ncell[2];
  int32_t use_gravity, use_decay;
  int32_t n, n_walls, n_hull;  /* n_walls, n_hull <= 255 */
  int32_t cand_slots, cell_cap, nblocks;
  double seg_box[8][4];        /* written by tdm_disk_setup: bounding boxes (min x,
                                  min y, max x, max y) of the first 8 segments */
} tdm_disk_params;

typedef struct tdm_disk_arrays {
  double *x, *v, *omega;       /* (N,2), (N,2), (N,): state at parity 0 */
  double *x_alt, *v_alt, *omega_alt;
  double *theta;               /* (N,) */
  const double *radius, *mass, *inertia;
  const double *walls, *hull;  /* (n_walls,2,2), (n_hull,2,2) segments */
  const double *wall_velocity; /* (n_walls,2) or NULL */
  double *out;                 /* TDM_DISK_NOUT */
  double *records;             /* NULL, or (cand_slots+n_walls+n_hull, N, TDM_DISK_NREC) */
  void *workspace;
  size_t workspace_bytes;
  int64_t launches, steps;     /* kept by the library: evaluations and steps issued so far */
} tdm_disk_arrays;

/* sizeof(tdm_disk_params) and sizeof(tdm_disk_arrays) as compiled. */
int tdm_disk_abi_sizes(size_t *params_bytes, size_t *arrays_bytes);
/* Blocks (256 disks each) for n disks. */
int tdm_disk_blocks(int n);
size_t tdm_disk_workspace_bytes(int n, int ncells, int cell_cap, int cand_slots, int n_walls,
                                int n_hull);
/* Clears lists, cell tables and contact state, schedules the first
 * candidate search; replaces DiskSystem.__init__'s neighbour and history
 * setup (granular2d/engine.py:194-209, 683-685). */
int tdm_disk_setup(tdm_disk_arrays *a, tdm_disk_params *p, void *stream);
/* Schedules a candidate search (after the caller moved disks). */
int tdm_disk_rebin(tdm_disk_arrays *a, const tdm_disk_params *p, void *stream);
/* Re-derives the segment bounding boxes (device table and p->seg_box) after
 * the caller moved wall or hull segments in place; contact state and lists
 * are kept.  The reference reads DiskSystem.walls / hull afresh at every
 * evaluation (granular2d/engine.py:352-375, 411-412), so callers move plates
 * and gates between steps (granular_identifiability/targets.py:241-246,
 * segregation/bed2d.py:170-174).  Waits for the boxes on `stream`. */
int tdm_disk_segments(tdm_disk_arrays *a, tdm_disk_params *p, void *stream);
/* nsteps steps; state alternates x/v/omega <-> x_alt/v_alt/omega_alt from
 * parity 0.  flags: TDM_DISK_LEDGER accumulates the energy ledger,
 * TDM_DISK_RECORD writes per-contact records on the last step.  Replaces
 * DiskSystem.step and _evaluate(advance_history=True)
 * (granular2d/engine.py:399-627). */
#define TDM_DISK_LEDGER 1
#define TDM_DISK_RECORD 2
int tdm_disk_step(tdm_disk_arrays *a, const tdm_disk_params *p, int nsteps, int flags,
                  void *stream);
/* Evaluation at the parity-0 state without advancing histories or state
 * (records when a->records is set); replaces DiskSystem.contact_forces
 * (granular2d/engine.py:675-678). */
int tdm_disk_inspect(tdm_disk_arrays *a, const tdm_disk_params *p, void *stream);
/* As tdm_disk_step (nsteps even, no flags) with HIP events on each step's
 * two kernels; mean durations in microseconds.  Measurement helper. */
int tdm_disk_step_timed(tdm_disk_arrays *a, const tdm_disk_params *p, int nsteps,
                        void *stream, double *mean_step_us, double *mean_prep_us);
/* 2*pairs steps captured as a graph (state ends in x/v/omega).  The graph
 * bakes in the parities of a->launches and a->steps at capture: replay it
 * only when both have those parities again, and add 2*pairs to both per
 * replay. */
int tdm_disk_graph_create(const tdm_disk_arrays *a, const tdm_disk_params *p, int pairs,
                          int flags, void *stream, tdm_graph **out);

/* ---------------------------------------------------------------------------
 * Bonded spheres in 3-D: the force evaluation and step of IceDEM3D
 * (src/tensordem/dem3d.py:564-1967), FP64, one simulation per call.
 * Options of the reference's DEM3DConfig outside these parameters (rolling
 * resistance, friction regularisation, Mindlin partial slip, platen
 * friction, buoyancy, Maxwell-Wiechert bonds, creep) are not evaluated here;
 * the host refuses them.
 * ------------------------------------------------------------------------- */
/* `out` (float64), written by every evaluation */
#define TDM_D3_REACTION 0       /* x, y, z: tool reaction, -sum of the tool force [N] */
#define TDM_D3_TOP_REACTION 3   /* top platen reaction z [N] */
#define TDM_D3_BOTTOM_REACTION 4
#define TDM_D3_LATERAL 5        /* x_min, x_max, y_min, y_max platen reactions [N] */
#define TDM_D3_FRICTION_POWER 9 /* sum of F_t . (v_i - v_j) over contacts [W] */
#define TDM_D3_CONTACTS 10      /* contacts evaluated (pairs without a live bond) */
#define TDM_D3_BREAKS 11        /* bonds broken by this evaluation */
#define TDM_D3_STATUS 12        /* TDM_ST_* bits, sticky until tdm_d3_clear_status */
#define TDM_D3_REBUILDS 13      /* candidate-list rebuilds so far */
#define TDM_D3_NOUT 14
/* flags of tdm_d3_forces */
#define TDM_D3_UPDATE_FRACTURE 1
#define TDM_D3_UPDATE_HISTORY 2

typedef struct tdm_d3_params {
  double radius, mass, dt, drag;
  double kb, kc, kt, mu;          /* bond, contact, tangential stiffness, friction */
  double eps_t, eps_s;            /* tensile (when thr is NULL) and shear thresholds */
  double hz_pair;                 /* (4/3) E* sqrt(R*), pair */
  double hz_pair_u;               /* (8/15) E* sqrt(R*), pair potential */
  double hz_rigid;                /* (4/3) E* sqrt(R*), sphere on a rigid surface */
  double hz_kt;                   /* 8 G* sqrt(R*), Mindlin tangential stiffness */
  double zeta2;                   /* 2 zeta of the restitution coefficient */
  double gamma_mobile, gamma_pinned; /* linear dashpots at m/2 and m */
  double cdamp;                   /* contact_damping: tool, and the absolute pair dashpot */
  double tool_r;
  double platen_k, platen_c, top_z0, bot_z0, top_v, bot_v, lat_k, lat_c;
  double mg;                      /* mass * gravity */
  double cell_w;                  /* cell width, at least 2 r + skin */
  double skin;                    /* candidate lists hold partners within 2 r + skin */
  int32_t hertz, damp;            /* damp: 0 none, 1 restitution, 2 absolute */
  int32_t friction, mindlin_kt, top, bot, lateral, gravity;
  int32_t n, nbonds, ncells, cell_cap, kh, nblocks;  /* ncells a power of two */
} tdm_d3_params;

typedef struct tdm_d3_arrays {
  double *x, *v;                  /* (N,3): state at entry and at exit */
  double *x_alt, *v_alt;          /* (N,3) */
  const double *x0;               /* (N,3) initial positions */
  const uint8_t *fixed;           /* (N,) */
  const int32_t *nbr;             /* (N,26) partners above (ascending), then below, -1 absent */
  const int32_t *bond_of_slot;    /* (N,26) bond index of each partner */
  const int32_t *pairs;           /* (B,2) bonds (i < j), sorted by (i, j) */
  uint8_t *alive;                 /* (B,) */
  const double *thr;              /* (B,) per-bond tensile threshold, or NULL */
  uint8_t *failure_mode;          /* (B,) 1 tensile, 2 shear, 3 both */
  double *failure_time_s, *failure_extension_m, *failure_energy_J; /* (B,) */
  const double *loads;            /* (N,3) nodal loads, or NULL */
  double *force;                  /* (N,3) written by every evaluation but energy */
  double *pair_energy;            /* (N,) written by tdm_d3_energy */
  const double *tool;             /* (6,) start, velocity */
  double *clock;                  /* (10,) time, steps, lateral positions (4), velocities (4) */
  double *out;                    /* TDM_D3_NOUT */
  void *workspace;
  size_t workspace_bytes;
  int64_t evals, hist_flips, steps; /* kept by the library */
} tdm_d3_arrays;

int tdm_d3_abi_sizes(size_t *params_bytes, size_t *arrays_bytes);
/* Blocks (256 spheres each) for n spheres. */
int tdm_d3_blocks(int n);
size_t tdm_d3_workspace_bytes(int n, int nbonds, int ncells, int cell_cap, int kh);
/* Inverse reference covariance of every sphere, cleared cell tables,
 * friction springs and status; replaces IceDEM3D.__init__'s derived state
 * (dem3d.py:579-760). */
int tdm_d3_setup(tdm_d3_arrays *a, const tdm_d3_params *p, void *stream);
int tdm_d3_clear_status(tdm_d3_arrays *a, const tdm_d3_params *p, void *stream);
/* Rebuilds the candidate lists in the next evaluation (after the caller
 * moved spheres). */
int tdm_d3_rebin(tdm_d3_arrays *a, const tdm_d3_params *p, void *stream);
/* Drops every friction spring (a restart: the reference's state_dict does
 * not carry them, dem3d.py:1828-1852). */
int tdm_d3_clear_history(tdm_d3_arrays *a, const tdm_d3_params *p, void *stream);
/* nsteps steps of IceDEM3D.step (dem3d.py:1644-1662), each one force
 * evaluation with fracture and friction history committed. */
int tdm_d3_step(tdm_d3_arrays *a, const tdm_d3_params *p, int nsteps, void *stream);
/* IceDEM3D.forces(external_forces, update_fracture, update_history)
 * (dem3d.py:1300-1642): force (N,3) and out, state unchanged. */
int tdm_d3_forces(tdm_d3_arrays *a, const tdm_d3_params *p, int flags, void *stream);
/* Contact energy of every pair without a live bond, to its lower sphere
 * (dem3d.py:1748-1772). */
int tdm_d3_energy(tdm_d3_arrays *a, const tdm_d3_params *p, void *stream);
/* Objective shear strain of every sphere at x (dem3d.py:875-901). */
int tdm_d3_shear(tdm_d3_arrays *a, const tdm_d3_params *p, double *shear, void *stream);
/* 2*pairs steps captured as a graph.  The graph bakes in the parities of
 * a->evals and a->hist_flips at capture; a replay leaves both unchanged. */
int tdm_d3_graph_create(tdm_d3_arrays *a, const tdm_d3_params *p, int pairs, void *stream,
                        tdm_graph **out);
/* As tdm_d3_step (nsteps even) with HIP events around each evaluation's
 * strain, bond and step kernels; mean durations in microseconds.
 * Measurement helper. */
int tdm_d3_step_timed(tdm_d3_arrays *a, const tdm_d3_params *p, int nsteps, void *stream,
                      double *mean_step_us, double *mean_strain_us, double *mean_bond_us);

#ifdef __cplusplus
}
#endif
#endif /* TENSORDEM_AMD_H */
