// ReplicaBatch: many independent samples advanced together on one GPU
// (rumdx_batch_* in include/rumdx.h).  RUMD runs one Sample at a time; this
// class is for ensembles of small systems, where per-sample launches leave
// the GPU idle.  Each replica has its own cubic box, particles, Nose-Hoover
// thermostat and neighbour-list rebuild schedule; all share one pair
// potential (its parameter table and cut-off method) and one time step.
#pragma once

#include <map>
#include <vector>

#include "rumd_host.h"

class ReplicaBatch {
 public:
  explicit ReplicaBatch(unsigned num_types);
  ~ReplicaBatch();
  ReplicaBatch(const ReplicaBatch&) = delete;
  ReplicaBatch& operator=(const ReplicaBatch&) = delete;

  // Adds a replica: positions and velocities n x 3 (in [-L/2, L/2)), types n.
  // Returns its index.
  unsigned AddReplica(const std::vector<float>& positions, const std::vector<unsigned>& types,
                      const std::vector<float>& velocities, float box_length);
  void SetMass(unsigned type, float mass);
  // The pair potential's table and cut-off method, and the list skin.
  void SetPotential(const PairPotential& pot, float skin);
  void SetTimeStep(float dt);
  // Nose-Hoover thermostat of one replica (relaxation time tau <= 0: NVE).
  void SetThermostat(unsigned replica, float target_temperature, float tau);
  void SetMaxNumNbrs(unsigned m);
  // Lanes that share one particle's list in the force kernel (1, 2, 4, 8).
  void SetLanesPerParticle(unsigned tp);
  unsigned GetLanesPerParticle() const { return lanes_per_particle_; }
  // Force kernel with each replica staged whole in LDS (on by default where
  // the largest replica fits the device's per-workgroup LDS), or neighbours
  // gathered through L2; the forces are the same bits.
  void SetStagedForces(bool on);
  bool GetStagedForces() const { return staged_forces_; }
  bool UsesStagedForces() const { return desc_.force_work_dev != nullptr; }
  // Rebuild sort: one workgroup per rebuilding replica in LDS (on by default
  // where the largest replica fits), or four kernels over the whole batch;
  // the same order either way.
  void SetWorkgroupSort(bool on);
  bool GetWorkgroupSort() const { return workgroup_sort_; }
  bool UsesWorkgroupSort() const { return desc_.max_rep_cells != 0; }
  // Steps with the forces and the NVT update in one launch
  // (rumdx_batch_forces_nvt; on by default with staged forces, whose work
  // items are then whole replicas).  The same bits either way.
  void SetFusedSteps(bool on);
  bool GetFusedSteps() const { return fused_steps_; }
  bool UsesFusedSteps() const { return desc_.force_work_dev != nullptr && desc_.force_work_whole != 0; }
  // Device copies of the dynamic state (positions, velocities, images, slot
  // ids, forces, thermostats, flags); RestoreState puts them back and marks
  // every list stale, so trial runs leave no trace in the trajectory.
  void SaveState();
  void RestoreState();

  // Forces and energies of the current positions (lists updated first).
  void CalcF();
  // nsteps steps of CalcF + leapfrog, as HIP-graph blocks of `block` steps.
  void RunSteps(unsigned nsteps, unsigned block);
  // Waits for the stream; grows the lists after an overflow.
  void Synchronize();

  unsigned GetNumberOfReplicas() const { return static_cast<unsigned>(reps_.size()); }
  unsigned GetNumberOfParticles(unsigned replica) const;
  unsigned GetTotalParticles() const;
  unsigned GetNumberOfSlots() const { return nslots_; }
  std::vector<float> GetPositions(unsigned replica);   // n x 3
  std::vector<float> GetVelocities(unsigned replica);  // n x 3
  std::vector<float> GetForces(unsigned replica);      // n x 4: F, u_i
  // Per replica {potential energy, virial (sum w_i / 6), kinetic energy, sum m v^2}.
  std::vector<double> GetEnergies();
  std::vector<double> GetThermostatStates();  // Ps per replica
  std::vector<unsigned> GetNumberOfBuilds();
  unsigned GetMaxNumNbrs() const { return max_nbrs_; }
  float GetBoxLength(unsigned replica) const;
  // {target temperature, relaxation time} of a replica's thermostat
  std::vector<float> GetThermostat(unsigned replica) const;
  float GetTimeStep() const { return dt_; }
  float GetSkin() const { return skin_; }
  int GetCutoffMethod() const { return cutoff_method_; }
  const std::vector<float>& GetParamsTable() const { return table_; }
  // Average duration (ms) of one rumdx_batch_forces launch over `reps`
  // launches on the current lists, timed with HIP events on the batch's stream.
  float TimeForceKernel(unsigned reps);
  // {mean, longest} list length over the real particles of every replica.
  std::vector<double> GetNeighborListStats();
  // Neighbour list of one replica in replica-local ids: CSR (offsets n + 1, ids).
  void GetNeighborList(unsigned replica, std::vector<uint32_t>& offsets, std::vector<uint32_t>& ids);
  hipStream_t GetStream() const { return stream_; }

 private:
  struct Spec {
    std::vector<float> pos, vel;
    std::vector<unsigned> types;
    float L = 0.f, T = 1.f, tau = 0.f;
  };
  void Allocate();
  void Relayout();
  void Step();
  void InvalidateGraphs();
  rumdx_replica HostReplica(unsigned r) const;
  std::vector<float> Slots(const DeviceArray<float>& a, unsigned replica, unsigned comps);
  // Replica-local particle id held in each of the replica's slots (the spatial
  // sort of every rebuild moves particles between slots).
  std::vector<uint32_t> SlotIds(unsigned replica);

  unsigned ntypes_;
  std::vector<Spec> reps_;
  std::vector<float> mass_;
  std::vector<float> table_;
  int cutoff_method_ = 2;
  float skin_ = 0.3f, max_cutoff_ = 0.f, dt_ = 0.0025f;
  unsigned user_max_nbrs_ = 0, max_nbrs_ = 0, lanes_per_particle_ = 1;
  bool saved_ = false;
  bool allocated_ = false;
  bool staged_forces_ = true;
  bool workgroup_sort_ = true;
  bool fused_steps_ = true;
  void BuildForceWork();
  uint32_t max_rep_n_ = 0, max_rep_cells_ = 0;
  hipStream_t stream_ = nullptr;

  std::vector<uint32_t> slot0_, cell0_;
  std::vector<int32_t> nc_;
  uint32_t nslots_ = 0, ncells_ = 0;
  DeviceArray<float> r_, v_, f_, w_, im_, last_r_, params_, cut2_, t_r_, t_v_, t_im_;
  DeviceArray<uint32_t> block_rep_, flags_, nbl_, num_nbrs_, cell_count_, cell_start_, cell_n_, cell_tmp_,
      slot_cell_, idx_, t_idx_, fixed_, cellblock_rep_, force_work_;
  DeviceArray<rumdx_replica> rep_;
  DeviceArray<rumdx_nh_state> nh_, save_nh_;
  DeviceArray<float> save_r_, save_v_, save_f_, save_w_, save_im_;
  DeviceArray<uint32_t> save_idx_, save_flags_;
  DeviceArray<double> workspace_, energies_;
  rumdx_batch desc_{};
  std::map<unsigned, hipGraphExec_t> graphs_;
};
