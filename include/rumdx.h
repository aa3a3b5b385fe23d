/*
 * rumdx.h -- C ABI of the MI355X (gfx950) implementation of RUMD's per-timestep
 * hot path: LJ 12-6 pair forces, neighbour lists with cell sorting, the
 * Nose-Hoover leapfrog integrator and the FP64 block reductions.
 *
 * Conventions
 *  - Every entry point returns 0 on success and a negative RUMDX_E* code on
 *    failure; rumdx_last_error() then returns a thread-local message.  This
 *    replaces RUMD's RUMD_Error exceptions (include/rumd/RUMD_Error.h:6-31),
 *    which the C++ layer above re-raises.
 *  - Pointers named *_dev are device pointers; everything else is host memory.
 *  - Per-particle arrays use RUMD's float4 layout (include/rumd/ParticleData.h:41-62)
 *    and are indexed by storage slot (the cell-sorted order after a sorted
 *    rebuild); unsorted_index[slot] is the original particle id
 *    (src/ParticleData.cu:383-393).
 *  - `stream` is a hipStream_t (NULL = the default stream).  No entry point
 *    synchronises, allocates or frees memory; scratch comes from buffers the
 *    caller sized with the *_bytes() helpers.
 *  - The box is RUMD's float[8] {cubic, Lx, Ly, Lz, 1/Lx, 1/Ly, 1/Lz, shift}
 *    (include/rumd/SimulationBox.h:165-178), passed by value from host memory.
 */
#ifndef RUMDX_H
#define RUMDX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RUMDX_VERSION 1

#define RUMDX_OK 0
#define RUMDX_E_INVALID -1   /* bad argument (RUMD_Error) */
#define RUMDX_E_RESOURCE -2  /* launch resources exceeded (RUMD_cudaResourceError) */
#define RUMDX_E_HIP -3       /* HIP runtime error */

/* PairPotential::CutoffMethod (include/rumd/PairPotential.h:35) */
#define RUMDX_CUTOFF_NS 0
#define RUMDX_CUTOFF_SP 1
#define RUMDX_CUTOFF_SF 2

/* Number of floats per type pair in the parameter table (include/rumd/rumd_base.h:28). */
#define RUMDX_NUM_PARAM 12

/* Pair potential forms served by rumdx_pair_form (RUMD's Pot_* classes
 * outside the 12-6 family; parameter rows as each class's SetParams stores
 * them, include/rumd/PairPotential.h). */
#define RUMDX_POT_LJ_SMOOTH_2_4 1 /* Pot_LJ_12_6_smooth_2_4 (:218-256) */
#define RUMDX_POT_GLJ_M_N 2       /* Pot_gLJ_m_n (:264-304) */
#define RUMDX_POT_GAUSS 3         /* Pot_Gauss (:383-414) */
#define RUMDX_POT_BUCKINGHAM 4    /* Pot_Buckingham (:422-461) */
#define RUMDX_POT_IPL_N 5         /* Pot_IPL_n (:600-636) */
#define RUMDX_POT_YUKAWA 6        /* Pot_Yukawa (:795-840) */
#define RUMDX_POT_EXP 7           /* Pot_Exp (:1019-1052) */
#define RUMDX_POT_LJ_GAUSS 8      /* Pot_LJ_12_6_Gauss (:336-378) */
#define RUMDX_POT_IPL_12 9        /* Pot_IPL_12 (:469-505) */
#define RUMDX_POT_IPL_9 10        /* Pot_IPL_9 (:512-549) */
#define RUMDX_POT_IPL_18 11       /* Pot_IPL_18 (:558-593) */
#define RUMDX_POT_IPL6_LN 12      /* Pot_IPL6_LN (:644-680) */
#define RUMDX_POT_LJ_LN 13        /* Pot_LJ_LN (:689-733) */
#define RUMDX_POT_GIRIFALCO 14    /* Pot_Girifalco (:743-786) */
#define RUMDX_POT_DZUGUTOV 15     /* Pot_Dzugutov (:846-900) */
#define RUMDX_POT_FERMI_JAGLA 16  /* Pot_Fermi_Jagla (:910-961) */
#define RUMDX_POT_REP_SHOULDER 17 /* Pot_Repulsive_Shoulder (:969-1012) */
#define RUMDX_POT_SAAP 18         /* Pot_SAAP (:1060-1135) */
#define RUMDX_POT_SF_COULOMB 19   /* Pot_ShiftedForceCoulomb (:1143-1186) */
/* Largest number of particle types a launch accepts (thermostat mask is 32 bits,
 * src/IntegratorNVT.cu:157-174). */
#define RUMDX_MAX_TYPES 32

/* Slots of rumdx_nblist.flags_dev (device uint32[RUMDX_NB_FLAGS]). */
#define RUMDX_NB_FLAGS 16
#define RUMDX_FLAG_REBUILD 0   /* != 0: the next build call rebuilds (updateRequired[0]) */
#define RUMDX_FLAG_OVERFLOW 1  /* sticky: some list hit max_nbrs (updateRequired[1]) */
#define RUMDX_FLAG_BUILDS 2    /* number of completed builds */
#define RUMDX_FLAG_MAXLEN 3    /* longest list seen since the flag was cleared (sorted builds: see rumdx_nblist_build_stats) */
#define RUMDX_FLAG_DONE 4      /* internal completion counter, zero between calls */
#define RUMDX_FLAG_MAXHALO 5   /* largest tile halo (particles) of the last sorted build, after rumdx_nblist_build_stats */
#define RUMDX_FLAG_ERROR 6     /* nonzero: the last sorted build is unusable, RUMDX_BUILD_ERR_* bits */
#define RUMDX_FLAG_EXCL_PENDING 7 /* internal: set by a completed build, cleared once its lists' exclusions are applied */
#define RUMDX_FLAG_MAXTILE 8   /* most particles in one cell tile, last sorted build, after rumdx_nblist_build_stats */
#define RUMDX_FLAG_PRUNE 9     /* != 0: the next rumdx_nblist_prune rebuilds the inner lists */
#define RUMDX_FLAG_STATS 10    /* internal: set by a completed sorted build, cleared by rumdx_nblist_build_stats */
/* Bits of flags[RUMDX_FLAG_ERROR]. */
#define RUMDX_BUILD_ERR_HALO 1u     /* a tile halo held more than 65535 particles */
#define RUMDX_BUILD_ERR_MOVED 2u    /* rebuild_sorted: a particle moved more than one cell since the last build */
#define RUMDX_BUILD_ERR_WAIT 4u     /* rebuild_sorted: a tile's wait for the tiles before it timed out */
#define RUMDX_BUILD_ERR_TILE 8u     /* rebuild_sorted: a tile held more than RUMDX_RESORT_TILE_MAX particles */
/* Particles one tile may hold for rumdx_nblist_rebuild_sorted. */
#define RUMDX_RESORT_TILE_MAX 2048u

/* Device-resident particle arrays (ParticleData, include/rumd/ParticleData.h:41-62). */
typedef struct rumdx_particles {
  float* r_dev;     /* nvp x 4: x, y, z, type (int bits in the float) */
  float* v_dev;     /* nvp x 4: vx, vy, vz, 1/m */
  float* f_dev;     /* nvp x 4: fx, fy, fz, u_i (pair energy, already halved) */
  float* w_dev;     /* nvp x 4: w_mol, sxz, sxy, w_i (sum_j r_ij^2 s_ij) */
  float* im_dev;    /* nvp x 4: image counts stored as floats */
  float* sts_dev;   /* nvp x 4: sxx, syy, szz, syz; may be NULL when stresses are never requested */
  float* rlo_dev;   /* nvp x 4: low-order position parts (FP64 position bookkeeping); NULL = plain FP32 positions */
  uint32_t* unsorted_index_dev; /* nvp: original particle id of each slot */
  uint32_t n;       /* real particles */
  uint32_t nvp;     /* allocated (padded) particles, >= n */
} rumdx_particles;

/* Neighbour-list state (NeighborList, include/rumd/NeighborList.h).  Two
 * entry formats:
 *   N2 method:     nbl_dev, u32 slot of the neighbour, column-major like
 *                  RUMD's: entry k of slot i at [k*nvp + i]
 *                  (src/PairPotential.cu:360-367);
 *   sorted method: nbl16_dev, u16 position of the neighbour inside the halo
 *                  of the particle's cell tile (see rumdx_tile_table): the
 *                  LDS byte offset 16*h of halo particle h when the halo holds
 *                  at most 4096 particles, else h itself; in groups of 8
 *                  consecutive entries: entry k of slot i at
 *                  [((k/8)*nvp + i)*8 + k%8], so one 16-byte load serves 8
 *                  pairs and a wave reads 1 KiB contiguous per group.  The
 *                  halo layout is rebuilt from the cell arrays, and
 *                  rumdx_nblist_export() converts to slots.  lane_order_dev
 *                  holds, per tile, the tile's particles (local index) by
 *                  descending list length, so the force kernel gives each
 *                  wave lists of similar length.  tile_halo_dev and
 *                  tile_info_dev record each tile's halo (slots in halo order)
 *                  as the build saw it, so the force kernel stages the halo
 *                  without rebuilding its table; they are sized by halo_cap,
 *                  and a change of halo_cap needs a rebuild before the next
 *                  force call (rumdx_nblist_request_rebuild).
 * num_nbrs_dev[i] == max_nbrs marks a list that overflowed; the force kernel
 * then rescans the cell stencil for that particle. */
typedef struct rumdx_nblist {
  uint32_t* nbl_dev;        /* N2 method: max_nbrs x nvp */
  uint32_t* num_nbrs_dev;   /* nvp */
  float* last_r_dev;        /* nvp x 4: positions at the last build */
  uint32_t* flags_dev;      /* RUMDX_NB_FLAGS x uint32, see RUMDX_FLAG_* */
  const float* cut_skin2_dev; /* ntypes^2: (rc_ij + skin)^2 formed in FP32 (src/NeighborList.cu:1239-1240) */
  uint32_t max_nbrs;
  uint32_t ntypes;
  float skin;
  /* Cell machinery for the sorted method (NULL/0 for the N2 method). */
  int32_t nc[3];            /* cells per axis, from rumdx_cell_grid() */
  const uint32_t* cell_rank_dev; /* nc0*nc1*nc2: linear cell -> storage rank (Morton order) */
  uint32_t* cell_start_dev; /* ncells + 1, by rank: first slot of each cell */
  uint32_t* cell_count_dev; /* ncells, by rank: zero between builds */
  uint32_t* slot_cell_dev;  /* nvp: linear cell index of each slot at the last build */
  void* sort_scratch_dev;   /* rumdx_sort_scratch_bytes(nvp, ncells) bytes, zero-filled before the first build */
  uint16_t* nbl16_dev;      /* sorted method: (max_nbrs rounded up to 8) x nvp halo offsets */
  uint32_t ntiles;          /* from rumdx_tile_table() */
  const uint32_t* tile_first_rank_dev; /* ntiles + 1 */
  const uint32_t* tile_origin_dev;     /* ntiles: cx0 | cy0 << 10 | cz0 << 20 */
  uint32_t halo_cap;        /* halo particles a tile stages in LDS (<= RUMDX_HALO_CAP_MAX) */
  uint16_t* lane_order_dev; /* sorted method: nvp, written by the build */
  uint32_t* tile_halo_dev;  /* sorted method: ntiles x halo_cap slots of each tile's halo, written by the build */
  uint32_t* tile_info_dev;  /* sorted method: ntiles x 4, written by the build */
  uint32_t lanes_per_particle; /* 1, 2, 4 or 8 lanes share a particle's list in the force
                                  kernels (RUMD's threads per particle, tp); 0 = 1 */
  /* Exclusions (NeighborList::SetExclusion, src/NeighborList.cu:791-933): the
   * particles each particle does not interact with through this list, CSR
   * over original ids (p->unsorted_index_dev maps slots to them); NULL = none.
   * The N2 build skips them, the sorted build removes them from its lists
   * after a rebuild (apply_exclusion_list, src/NeighborList.cu:1323-1432), and
   * the force kernels' cell rescans skip them.  NB method "none" is not
   * available with exclusions (the reference switches to N2). */
  const uint32_t* excl_start_dev; /* n + 1 */
  const uint32_t* excl_ids_dev;
  /* Inner lists of the sorted method (NULL = none; no counterpart in the
   * reference).  The tile list holds every pair within rc_ij + skin at the
   * last build; the inner list of slot i holds the entries of that list whose
   * pair was within rc_ij + inner_skin at the last prune, in the same order
   * and format (nbl16_in_dev like nbl16_dev, num_in_dev like num_nbrs_dev).
   * It stays valid while every particle moves less than inner_skin/2 since
   * the prune (the skin check raises flags[RUMDX_FLAG_PRUNE] at 0.49
   * inner_skin), and the force kernels read it instead of the tile list, so
   * pairs in the skin shell are not evaluated on every step. */
  uint16_t* nbl16_in_dev;   /* like nbl16_dev */
  uint32_t* num_in_dev;     /* nvp */
  float* prune_r_dev;       /* nvp x 4: positions at the last prune */
  const float* cut_in2_dev; /* ntypes^2: (rc_ij + inner_skin)^2 in FP32 */
  float inner_skin;
  /* Occupancy the 12-6 tile force kernel is compiled for, in waves per SIMD:
   * 0 = the default (5), or 4 or 6 (compiled variants of the stepping
   * kernel, for the autotuner).  Results are bit-identical. */
  uint32_t tile_waves_per_eu;
} rumdx_nblist;

/* Cells per tile edge and the largest LDS halo a launch may request. */
#define RUMDX_TILE_EDGE 4
#define RUMDX_HALO_CAP_MAX 3072
/* tile_info_dev[4 t + 1] of a sorted build: bit 0 set when the tile's halo
 * was staged and every pair takes the plain FP32 difference (in the shifted
 * frame where the halo wraps); bits 1-6 mark, per axis, a halo wrapping below
 * (bit 2a+1) or above (bit 2a+2) the tile; bits 16-31 hold the halo capacity
 * the build used.  tile_info_dev[4 t] is the tile's halo size, [4 t + 2] and
 * [4 t + 3] its slot range. */
#define RUMDX_TILE_FAST 1u
#define RUMDX_TILE_SHIFT_MASK 0x7Eu

/* Device-resident Nose-Hoover state (IntegratorNVT, src/IntegratorNVT.cu:82-92, 426-437). */
typedef struct rumdx_nh_state {
  double ps;          /* thermostat state Ps (d_thermostatState) */
  double target_T;    /* target temperature, advanced by dT_per_step after each update */
  double dT_per_step; /* SetTemperatureChangePerTimeStep */
  double ke2_sum;     /* FP64 sum of m v^2(t+h/2) over thermostatted particles, last step */
} rumdx_nh_state;

/* ---------------------------------------------------------------- misc */
int rumdx_version(void);
const char* rumdx_last_error(void);
/* Bytes of device scratch the reduction/integrator entry points need for n particles. */
size_t rumdx_workspace_bytes(uint32_t n);
/* Bytes of sort scratch a sorted build needs. */
size_t rumdx_sort_scratch_bytes(uint32_t nvp, uint32_t ncells);

/* ---------------------------------------------------------------- cells
 * Cell grid of NeighborList::CalculateNBL_Sorting (src/NeighborList.cu:717-726):
 * nc = max(5, floor(2 L / (max_cutoff + skin))) per axis, computed in FP32. */
int rumdx_cell_grid(const float box[8], float max_cutoff, float skin, int32_t nc_out[3]);
/* Storage rank of every linear cell (x fastest, src/NeighborList.cu:1136-1140):
 * cells are stored in Morton (Z-order) order of their coordinates.  Host output. */
int rumdx_cell_rank_table(const int32_t nc[3], uint32_t* rank_out);
/* Tiles: aligned RUMDX_TILE_EDGE^3 bricks of cells in Morton order, i.e.
 * contiguous rank (and slot) ranges.  Returns the tile count; with non-null
 * outputs also fills first_rank_out[ntiles+1] and origin_out[ntiles] (brick
 * origin cell, packed cx | cy << 10 | cz << 20).  rank is the table of
 * rumdx_cell_rank_table.  Host only; returns a negative code on error. */
int rumdx_tile_table(const int32_t nc[3], const uint32_t* rank, uint32_t* first_rank_out, uint32_t* origin_out);
/* Halo capacity (particles) to request for a grid of nc cells holding n
 * particles: the mean 8^3-cell halo plus 8 % and 64, rounded up to 64, at
 * most RUMDX_HALO_CAP_MAX.  The force kernel's LDS is 16 bytes per halo
 * particle, so a smaller capacity lets more tiles share a CU.  Tiles whose
 * halo is larger still compute exactly, reading positions from global
 * memory. */
uint32_t rumdx_halo_cap_estimate(const int32_t nc[3], uint32_t n);

/* ---------------------------------------------------------------- neighbour list
 * NeighborList::CheckRebuildRequired / displacement_since_nb_build
 * (src/NeighborList.cu:569-581, 970-997): sets flags[REBUILD] if any particle
 * moved more than skin/2 (FP32 minimum image, strict >).  With inner lists it
 * also sets flags[PRUNE] if any particle moved more than 0.49 inner_skin since
 * the last prune. */
int rumdx_nblist_check(const rumdx_particles* p, rumdx_nblist* nb, const float box[8], void* stream);

/* NeighborList::ResetNeighborList (src/NeighborList.cu:504-509): marks the list
 * stale so that the next build call rebuilds it (a one-thread kernel, so it is
 * safe inside captured graphs). */
int rumdx_nblist_request_rebuild(rumdx_nblist* nb, void* stream);

/* NeighborList::UpdateNBlist, N2 method (calculateNBL_N2, src/NeighborList.cu:1001-1099).
 * Does nothing on the device unless flags[REBUILD] is set. */
int rumdx_nblist_build_n2(const rumdx_particles* p, rumdx_nblist* nb, const float box[8], void* stream);

/* NeighborList::UpdateNBlist, sort method: cell binning, spatial sort of every
 * particle array and the cell-stencil build (CalculateNBL_Sorting,
 * src/NeighborList.cu:717-781, 1142-1319; Sample::SortParticlesByKey,
 * include/rumd/Sample.h:174-188; ParticleData::UpdateAfterSorting,
 * src/ParticleData.cu:493-528).  Particles are stored by cell in Morton order of
 * the cells and by original slot inside a cell (a stable sort).  Does nothing
 * on the device unless flags[REBUILD] is set. */
int rumdx_nblist_build_sorted(rumdx_particles* p, rumdx_nblist* nb, const float box[8], void* stream);

/* The host's read of the build results before it grows the list
 * (src/NeighborList.cu:660-670): for the sorted method, gathers the last
 * build's longest list into flags[MAXLEN] (the maximum with its old value),
 * its largest tile halo into flags[MAXHALO] and its most crowded tile into
 * flags[MAXTILE] from the per-tile records the build left.  A one-workgroup
 * kernel that does nothing unless a sorted build completed since the last
 * call; the build itself leaves this out, so that the steps between host reads
 * carry no launch for it.  Call it before reading those flags.  Returns
 * RUMDX_OK without a launch when the list has no sorted build arrays. */
int rumdx_nblist_build_stats(const rumdx_particles* p, rumdx_nblist* nb, void* stream);

/* The same result as rumdx_nblist_build_sorted -- particle order, cells and
 * lists bit for bit -- for particles that a previous sorted build of this list
 * left in cell order and that have since moved less than one cell (the state
 * between skin-triggered rebuilds; the skin check keeps displacements below
 * skin/2 + one step, well under a cell of (rc + skin)/2).  Each cell tile
 * re-sorts only the particles of its own cells and their neighbours (no global
 * histogram or scatter), and tiles find their first slot from the counts of
 * the tiles before them.  Three launches, each returning at once unless a
 * rebuild is pending.  A broken precondition is reported in
 * flags[RUMDX_FLAG_ERROR] (RUMDX_BUILD_ERR_MOVED, _TILE), never as an out-of-
 * range access.  The build of NeighborList::UpdateNBlist for skin-triggered
 * rebuilds (src/NeighborList.cu:583-690, 717-781); no separate counterpart. */
int rumdx_nblist_rebuild_sorted(rumdx_particles* p, rumdx_nblist* nb, const float box[8], void* stream);

/* Prunes the tile lists into the inner lists (nb->nbl16_in_dev): for every
 * particle of a tile whose halo the build staged, the entries whose pair is
 * within rc_ij + inner_skin at the current positions, in list order (FP32
 * distance; the displacement is exact in the tile's frame or RUMD's FP32
 * minimum image), and records the positions in prune_r_dev.  Other tiles'
 * particles and overflowed lists keep the tile list's count (the force
 * kernels rescan them).  Runs on the device only when flags[RUMDX_FLAG_PRUNE]
 * is set -- by every sorted build and by the skin check -- and clears it.
 * Call after the build and before the force call.  No counterpart in the
 * reference. */
int rumdx_nblist_prune(const rumdx_particles* p, rumdx_nblist* nb, const float box[8], void* stream);

/* Writes the current list in RUMD's format -- u32 neighbour slots, column
 * major, out_dev[k*nvp + i] for k < num_nbrs[i] -- whatever the internal
 * format (for tools and checks that read lists; not on the step path). */
int rumdx_nblist_export(const rumdx_particles* p, const rumdx_nblist* nb, uint32_t* out_dev, void* stream);

/* ---------------------------------------------------------------- pair forces
 * Pot_LJ_12_6 / Pot_IPL_12_6 force, energy and virial (PairPotential::CalcF,
 * kernels Calcf_NBL_tp_equal_one / Calcf_NBL_tp, src/PairPotential.cu:330-515;
 * fij, include/rumd/PairPotential.h:1198-1244).  params_dev is RUMD's table
 * ntypes*ntypes*12 floats (param 0 = Rcut, 1 = A12, 2 = A6, 10 = s(rc)*rc,
 * 11 = u(rc); src/PairPotential.cu:82-125).  nb == NULL sums over all pairs
 * (NB method "none").  initialize != 0 stores f/w (and sts), otherwise adds.
 * Pair arithmetic is FP32; the minimum-image displacement is formed in FP64. */
int rumdx_pair_lj(const rumdx_particles* p, const rumdx_nblist* nb, const float* params_dev,
                  uint32_t ntypes, const float box[8], int cutoff_method, int initialize,
                  int calc_stresses, void* stream);

/* All-FP64 variant of rumdx_pair_lj, the comparison path of the precision
 * study: positions r (+ rlo when set) widened to double, the minimum image,
 * the LJ term, the cut-off shifts (recomputed in double from rc, A12, A6) and
 * the per-particle sums all in double, rounded once into the float4 outputs.
 * Same arguments, lists and error behaviour as rumdx_pair_lj.  No
 * counterpart in the reference, whose precision is a build option. */
int rumdx_pair_lj_fp64(const rumdx_particles* p, const rumdx_nblist* nb, const float* params_dev,
                       uint32_t ntypes, const float box[8], int cutoff_method, int initialize,
                       int calc_stresses, void* stream);

/* Force, energy and virial of another pair potential form (RUMDX_POT_*):
 * PairPotential::CalcF of that Pot_* class, whose kernels the reference
 * generates from one skeleton per potential (src/Generate_PP_FunctionBodies.py:3-127,
 * ComputeInteraction of each class in include/rumd/PairPotential.h).
 * params_dev is RUMD's table as for rumdx_pair_lj (param 0 = Rcut, 1..9 the
 * form's parameters, 10 = s(rc)*rc, 11 = u(rc)).  Same lists, cut-off
 * methods, stress option, initialize semantics and errors as rumdx_pair_lj;
 * pair arithmetic in FP32 with FP32 exp/log/sqrt. */
int rumdx_pair_form(const rumdx_particles* p, const rumdx_nblist* nb, const float* params_dev,
                    uint32_t ntypes, const float box[8], int form, int cutoff_method, int initialize,
                    int calc_stresses, void* stream);

/* ---------------------------------------------------------------- intramolecular forces
 * Topology lists are CSR over original particle ids (what MoleculeData holds,
 * src/MoleculeData.cu:479-700): particle o's entries are [start[o], start[o+1]).
 * Kernels reach particles through slot_of (original id -> current slot), which
 * rumdx_slot_index derives from p->unsorted_index_dev; refresh it after any
 * list build that sorts.  A thread owns one particle and gathers its share of
 * every term it belongs to (no atomics, reproducible sums); initialize != 0
 * stores the particle's bonded sums (zero for a particle with none),
 * otherwise adds them. */

/* slot_of[unsorted_index[s]] = s for s < n. */
int rumdx_slot_index(const rumdx_particles* p, uint32_t* slot_of_dev, void* stream);

/* BondHarmonic::CalcF (src/BondPotential.cu:70-135, kernel Harmonic :281-357):
 * u = k/2 (r - l0)^2 per bond; each end gets its force, half the energy in
 * f.w and the virial ft r^2 in w.w (and the stress in sts, w.y, w.z when
 * calc_stresses).  bond_other/bond_type hold, for each entry, the other end's
 * original id and the bond type; params_dev = (l0, k) per type. */
int rumdx_bond_harmonic(const rumdx_particles* p, const uint32_t* slot_of_dev,
                        const uint32_t* bond_start_dev, const uint32_t* bond_other_dev,
                        const uint32_t* bond_type_dev, const float* params_dev, uint32_t ntypes,
                        const float box[8], int initialize, int calc_stresses, void* stream);

/* BondFENE::CalcF (kernel FENE, src/BondPotential.cu:358-430): u = -k R0^2/2
 * ln(1 - r^2/R0^2) per bond, params_dev = (R0, k) per type; otherwise as
 * rumdx_bond_harmonic. */
int rumdx_bond_fene(const rumdx_particles* p, const uint32_t* slot_of_dev, const uint32_t* bond_start_dev,
                    const uint32_t* bond_other_dev, const uint32_t* bond_type_dev, const float* params_dev,
                    uint32_t ntypes, const float box[8], int initialize, int calc_stresses, void* stream);

/* DihedralRyckaert::CalcF (kernel Ryckaert, src/DihedralPotential.cu:83-110,
 * 150-260): u = sum_k c_k cos^k(psi), psi = 0 for trans, per dihedral
 * (a, b, c, d).  dihedral_atoms_dev holds 5 u32 per entry (a, b, c, d,
 * type), listed under each of its four atoms; params_dev = c_0..c_5 per type.
 * The energy goes to atom a's f.w; no virial or stress is added (with
 * initialize, w and sts are cleared as for angles). */
int rumdx_dihedral_ryckaert(const rumdx_particles* p, const uint32_t* slot_of_dev,
                            const uint32_t* dihedral_start_dev, const uint32_t* dihedral_atoms_dev,
                            const float* params_dev, uint32_t ntypes, const float box[8], int initialize,
                            int calc_stresses, void* stream);

/* PeriodicDihedral::CalcF (src/DihedralPotential.cu:270-300, kernel PeriodicDih
 * :337-458): params (psi0, k1, k2, k3, 0, 0) per type, psi = acos(cos psi) in
 * [0, pi], u = 1/2 (k1 (1 - cos x) + k2 (1 - cos 2x) + k3 (1 - cos 3x)) with
 * x = psi - psi0, and the reference's forces, -du/dpsi times the gradient of
 * cos psi (-sin psi times the gradient force).  Lists, energy placement and
 * errors as rumdx_dihedral_ryckaert. */
int rumdx_dihedral_periodic(const rumdx_particles* p, const uint32_t* slot_of_dev,
                            const uint32_t* dihedral_start_dev, const uint32_t* dihedral_atoms_dev,
                            const float* params_dev, uint32_t ntypes, const float box[8], int initialize,
                            int calc_stresses, void* stream);

/* AngleCosSq::CalcF (src/AnglePotential.cu:69-98, kernel CosSq :137-217):
 * u = k/2 (cos theta - cos theta0)^2 per angle (a, b, c), the angle at b.
 * angle_atoms_dev holds 4 u32 per entry: a, b, c (original ids) and the angle
 * type, listed under each of its three atoms; params_dev = (theta0, k) per
 * type.  The angle's energy goes to atom a's f.w and no virial or stress is
 * added, as in the reference; with initialize, w (and sts when calc_stresses)
 * are cleared for the potentials that add after it. */
int rumdx_angle_cossq(const rumdx_particles* p, const uint32_t* slot_of_dev,
                      const uint32_t* angle_start_dev, const uint32_t* angle_atoms_dev,
                      const float* params_dev, uint32_t ntypes, const float box[8], int initialize,
                      int calc_stresses, void* stream);

/* AngleSq::CalcF (src/AnglePotential.cu:220-250, kernel SquaredAngle :289-373):
 * u = k/2 (theta - theta0)^2 per angle (a, b, c), theta the angle at b, with
 * the reference's force factor -k (theta - theta0) / (1e-6 + sin theta).
 * Lists, parameters, energy placement and errors as rumdx_angle_cossq. */
int rumdx_angle_sq(const rumdx_particles* p, const uint32_t* slot_of_dev, const uint32_t* angle_start_dev,
                   const uint32_t* angle_atoms_dev, const float* params_dev, uint32_t ntypes, const float box[8],
                   int initialize, int calc_stresses, void* stream);

/* ---------------------------------------------------------------- integrator
 * IntegratorNVT::Integrate (integrateNVTAlgorithm, src/IntegratorNVT.cu:64-79,
 * 364-418) followed, when thermostat_on_mask != 0, by UpdateThermostatState
 * (src/IntegratorNVT.cu:82-92, 426-437) with the kinetic-energy sum in FP64 and
 * Ps, T in FP64.  omega2 = 4 pi^2 / tau^2, dof = degrees of freedom.  When
 * p->rlo_dev != NULL positions are advanced and wrapped in FP64 and kept as
 * hi/lo float pairs.  When nb != NULL the skin check of rumdx_nblist_check is
 * fused in (flags[REBUILD] set for the next build). */
int rumdx_nvt_step(rumdx_particles* p, rumdx_nh_state* nh_dev, float dt, uint32_t thermostat_on_mask,
                   float omega2, uint32_t dof, const float box[8], rumdx_nblist* nb,
                   void* workspace_dev, void* stream);

/* One step of the run loop (Simulation.Run, Python/rumd/Simulation.py:583-597:
 * sample.CalcF then itg.Integrate) for a sample whose only force term is a
 * 12-6 pair potential over a list of the sorted method: rumdx_pair_lj with
 * initialize = 1 and no stresses, and rumdx_nvt_step with the skin check, in
 * one kernel.  A cell tile's particles are advanced as soon as the tile's
 * forces are complete; their new positions go to r_out_dev (nvp x 4, not p->r_dev: other particles'
 * forces still read the old positions), while v, im and rlo are updated in
 * place and f, w hold the forces, energies and virials at the old positions.
 * The caller swaps p->r_dev and r_out_dev afterwards.  With the thermostat on,
 * the kinetic-energy sum is taken per cell tile (one double per tile in the
 * workspace of rumdx_workspace_bytes(n), which must hold nb->ntiles of them:
 * RUMDX_E_INVALID otherwise) and then in tile order, so Ps agrees with
 * rumdx_nvt_step's to rounding; forces, velocities and positions of a step
 * are bit-identical to the two calls'. */
int rumdx_pair_lj_nvt(const rumdx_particles* p, const rumdx_nblist* nb, const float* params_dev, uint32_t ntypes,
                      const float box[8], int cutoff_method, float* r_out_dev, rumdx_nh_state* nh_dev, float dt,
                      uint32_t thermostat_on_mask, float omega2, uint32_t dof, void* workspace_dev, void* stream);

/* IntegratorNVT::SetMomentumToZero (src/IntegratorNVT.cu:176-183, 444-472):
 * subtracts the centre-of-mass velocity, with the momentum sum in FP64. */
int rumdx_zero_momentum(rumdx_particles* p, void* workspace_dev, void* stream);

/* IntegratorNVT::GetKineticEnergy (src/IntegratorNVT.cu:316-356): the leapfrog
 * average 1/2 sum m (v^2(t-h/2) + v^2(t+h/2))/2 with v(t+h/2) predicted from the
 * current forces and Ps.  Writes one double to out_dev. */
int rumdx_kinetic_energy(const rumdx_particles* p, const rumdx_nh_state* nh_dev, float dt,
                         uint32_t thermostat_on_mask, double* out_dev, void* workspace_dev,
                         void* stream);

/* ---------------------------------------------------------------- reductions
 * FP64 sums of float4 components (the sums behind Sample::GetPotentialEnergy,
 * GetVirial, src/Sample.cu:353-391, and sumIdenticalArrays,
 * src/rumd_algorithms.cu:162-250).  out_dev[c] = sum_i arr[i].c for the
 * components whose bit is set in comp_mask (bit 0 = x ... bit 3 = w); other
 * outputs are written as 0.  The summation order is fixed, so results are
 * bitwise reproducible. */
int rumdx_reduce_f4(const float* arr4_dev, uint32_t n, uint32_t comp_mask, double* out_dev,
                    void* workspace_dev, void* stream);

/* rumdx_reduce_f4 without a caller workspace, for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  No counterpart in the reference.  The
 * block partials live in a file-scope device array of the library (1 MiB,
 * one copy per GPU, held until the process ends), so n is limited to
 * RUMDX_STATIC_REDUCE_MAX_N (larger n: RUMDX_E_INVALID).  The same kernels
 * and summation order as rumdx_reduce_f4, so the results are bit-identical.
 * Calls that overlap on one GPU (different streams or threads) share the
 * array and race; serialise them.  Allocates nothing. */
#define RUMDX_STATIC_REDUCE_MAX_N 8388608u
int rumdx_reduce_f4_static(const float* arr4_dev, uint32_t n, uint32_t comp_mask, double* out_dev, void* stream);

/* ---------------------------------------------------------------- replicas
 * Batched independent samples ("replicas") in one set of device arrays,
 * advanced by one launch per kernel for all of them.  No counterpart in the
 * reference, which runs one Sample at a time; it is for workloads of many
 * small systems (parameter sweeps, ensembles), where per-sample launches
 * leave the GPU idle.  Each replica follows the single-sample algorithms
 * above: a cubic box, RUMD's cell grid and 5^3 stencil with the FP32
 * criterion, u32 lists (entry k of slot i at [k*nslots + i]), the LJ pair
 * term in FP32 with exact minimum-image displacements (taken on fixed-point
 * copies of the positions, fixed_dev), and its own Nose-Hoover leapfrog.
 * Replica r owns slots [slot0, slot0 + n); slot0 is a multiple of
 * RUMDX_BATCH_ALIGN and the range is padded to ceil(n / RUMDX_BATCH_ALIGN)
 * blocks, so no workgroup spans two replicas and each replica's sums have a
 * fixed order: a replica's trajectory does not depend on the others.  Each
 * rebuild sorts the replica's particles by cell inside its slot range (a
 * stable counting sort); idx_dev follows the particles, so idx_dev[slot]
 * names the replica-local particle a slot holds. */
#define RUMDX_BATCH_ALIGN 256
#define RUMDX_BATCH_FLAGS 8  /* u32 per replica: rebuild, overflow, builds, longest list, internal */
#define RUMDX_BATCH_CELL_BLOCK 16  /* a replica's cell range starts at a multiple of this and is padded to it */

typedef struct rumdx_replica {
  uint32_t slot0;  /* first slot, a multiple of RUMDX_BATCH_ALIGN */
  uint32_t n;      /* particles */
  uint32_t cell0;  /* first cell of its grid in the batch's cell arrays, a multiple of RUMDX_BATCH_CELL_BLOCK */
  int32_t nc;      /* cells per axis, rumdx_cell_grid of its box */
  float L;         /* cubic box length */
  float inv_L;     /* 1/L rounded to float, as RUMD's box[4] */
  float omega2;    /* 4 pi^2 / tau^2 of its thermostat; 0 = NVE */
  uint32_t dof;    /* degrees of freedom, 3n - 3 */
} rumdx_replica;

typedef struct rumdx_batch {
  float* r_dev;    /* nslots x 4, layout of rumdx_particles */
  float* v_dev;
  float* f_dev;
  float* w_dev;
  float* im_dev;
  uint32_t nslots; /* sum of the padded replica ranges */
  uint32_t nrep;
  const rumdx_replica* rep_dev;  /* nrep */
  const uint32_t* block_rep_dev; /* nslots / RUMDX_BATCH_ALIGN: replica of each slot block */
  rumdx_nh_state* nh_dev;        /* nrep */
  uint32_t* rep_flags_dev;       /* nrep x RUMDX_BATCH_FLAGS */
  uint32_t* nbl_dev;             /* max_nbrs x nslots */
  uint32_t* num_nbrs_dev;        /* nslots; == max_nbrs marks an overflowed list */
  float* last_r_dev;             /* nslots x 4 */
  const float* cut_skin2_dev;    /* ntypes^2: (rc_ij + skin)^2 in FP32 */
  uint32_t max_nbrs;
  uint32_t ntypes;
  float skin;
  uint32_t ncells;               /* sum over replicas of nc^3, each padded to RUMDX_BATCH_CELL_BLOCK */
  const uint32_t* cellblock_rep_dev; /* ncells / RUMDX_BATCH_CELL_BLOCK: replica of each cell block */
  uint32_t* cell_count_dev;      /* ncells, zero between builds */
  uint32_t* cell_start_dev;      /* ncells: first slot of each cell (cells are contiguous slot ranges) */
  uint32_t* cell_n_dev;          /* ncells */
  uint32_t* cell_tmp_dev;        /* nslots */
  uint32_t* slot_cell_dev;       /* nslots: cell of each slot at its replica's last build */
  uint32_t* idx_dev;             /* nslots: replica-local id of the particle in each slot */
  uint32_t lanes_per_particle;   /* 0/1, 2, 4 or 8 lanes share a list in rumdx_batch_forces */
  int32_t* fixed_dev;            /* nslots x 4: (x/L) 2^32 per axis as int32 and the type bits; written by
                                    rumdx_batch_update_lists and rumdx_batch_nvt_step, read by the forces */
  float* t_r_dev;                /* nslots x 4: sort staging of r, v, im and idx */
  float* t_v_dev;
  float* t_im_dev;
  uint32_t* t_idx_dev;
  /* Replica-staged forces (optional; NULL: neighbours gathered through L2).
   * Work item k is (replica, first slot, end slot, 0), a slot range of one
   * replica: rumdx_batch_forces runs one 1024-thread workgroup per item, which
   * copies the replica's fixed-point positions into LDS (max_rep_n x 16 bytes
   * plus the parameter table: at most the device's per-workgroup opt-in LDS)
   * and reads every neighbour from there.  The forces are the same bits. */
  const uint32_t* force_work_dev; /* force_work_n x 4 */
  uint32_t force_work_n;
  uint32_t max_rep_n;            /* particles of the largest replica */
  uint32_t force_work_whole;     /* nonzero: every work item is one whole replica (rumdx_batch_forces_nvt needs it) */
  /* Nonzero: rumdx_batch_update_lists sorts each rebuilding replica with one
   * workgroup in LDS (its cell counts, their scan and the stable placement),
   * which needs 4 (max_rep_n + 2 max_rep_cells) bytes of LDS, at most 64 KiB;
   * the largest nc^3 of any replica.  0: the four-kernel sort over the batch.
   * Both give the same order. */
  uint32_t max_rep_cells;
} rumdx_batch;

/* rumdx_batch_forces followed by rumdx_batch_nvt_step in one launch (and the
 * thermostat's finishing kernel): with a staged work list of whole replicas
 * (force_work_whole), each workgroup computes its replica's forces and then
 * advances it in place, with the same per-256-slot kinetic-energy partials as
 * rumdx_batch_nvt_step; the results are the same bits as the two calls'.  No
 * counterpart in the reference. */
int rumdx_batch_forces_nvt(rumdx_batch* b, const float* params_dev, int cutoff_method, float dt,
                           void* workspace_dev, void* stream);
/* Bytes of device workspace rumdx_batch_nvt_step / rumdx_batch_energies need. */
size_t rumdx_batch_workspace_bytes(uint32_t nslots);
/* Marks every replica's list stale. */
int rumdx_batch_request_rebuild(rumdx_batch* b, void* stream);
/* Rebuilds the lists of the replicas whose rebuild flag is set (the skin
 * check of rumdx_batch_nvt_step sets it); the others are untouched. */
int rumdx_batch_update_lists(rumdx_batch* b, void* stream);
/* LJ 12-6 forces, energies (f.w) and virials (w.w) of every replica, one
 * shared parameter table (rumdx_pair_lj's), stored (initialize semantics).
 * A particle whose list overflowed rescans its build-time cell stencil, so
 * forces stay exact; the host regrows max_nbrs at its next sync point. */
int rumdx_batch_forces(rumdx_batch* b, const float* params_dev, int cutoff_method, void* stream);
/* One leapfrog step of every replica (rumdx_nvt_step's update, all types
 * thermostatted when omega2 > 0), its thermostat update, and the skin check
 * that flags the replica's next rebuild. */
int rumdx_batch_nvt_step(rumdx_batch* b, float dt, void* workspace_dev, void* stream);
/* Per replica, out_dev[4r..4r+3] = {sum u_i, sum w_i, leapfrog kinetic
 * energy (as rumdx_kinetic_energy), sum m v^2}, FP64 in a fixed order. */
int rumdx_batch_energies(rumdx_batch* b, float dt, double* out_dev, void* workspace_dev, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* RUMDX_H */
