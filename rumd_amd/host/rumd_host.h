// Host-side C++ classes mirroring RUMD's Sample / Potential / Integrator API
// (include/rumd/Sample.h, Potential.h, PairPotential.h, NeighborList.h,
// Integrator.h, IntegratorNVT.h, SimulationBox.h, ParticleData.h,
// RUMD_Error.h, Device.h).  Every computation on particles runs on the GPU
// through the C ABI in include/rumdx.h; this layer owns device memory, the
// stream, parameter tables and the step schedule.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "rumdx.h"

// ------------------------------------------------------------------ errors
// RUMD_Error (include/rumd/RUMD_Error.h:6-31): class, method and message.
class RUMD_Error : public std::runtime_error {
 public:
  RUMD_Error(const std::string& className, const std::string& methodName, const std::string& errorStr)
      : std::runtime_error(className + "::" + methodName + ": " + errorStr),
        className(className),
        methodName(methodName),
        errorStr(errorStr) {}
  std::string className, methodName, errorStr;
};

// Resource errors map to ValueError in Python (Swig/rumd.i:41-53).
class RUMD_cudaResourceError : public RUMD_Error {
 public:
  using RUMD_Error::RUMD_Error;
};

void hip_check(hipError_t e, const char* cls, const char* method);
void rumdx_check(int rc, const char* cls, const char* method);

// ------------------------------------------------------------------ device memory
template <class T>
class DeviceArray {
 public:
  DeviceArray() = default;
  explicit DeviceArray(size_t n) { resize(n); }
  ~DeviceArray() { release(); }
  DeviceArray(const DeviceArray&) = delete;
  DeviceArray& operator=(const DeviceArray&) = delete;
  DeviceArray(DeviceArray&& o) noexcept : ptr_(o.ptr_), n_(o.n_) { o.ptr_ = nullptr; o.n_ = 0; }
  DeviceArray& operator=(DeviceArray&& o) noexcept {
    if (this != &o) {
      release();
      ptr_ = o.ptr_;
      n_ = o.n_;
      o.ptr_ = nullptr;
      o.n_ = 0;
    }
    return *this;
  }
  // Allocates n elements, zero-filled (synchronous; never on a hot path).
  void resize(size_t n);
  void release();
  T* get() const { return ptr_; }
  size_t size() const { return n_; }
  void upload(const T* h, size_t n, hipStream_t st);
  void download(T* h, size_t n, hipStream_t st) const;

 private:
  T* ptr_ = nullptr;
  size_t n_ = 0;
};

// ------------------------------------------------------------------ Device
// Device singleton (src/Device.cu:143-222): name, memory, sync, timing.
class Device {
 public:
  static Device& GetDevice();
  void Init();
  std::string GetDeviceName();
  std::string GetDeviceReport();
  size_t GetDeviceMemory();
  void Synchronize();
  float Time();  // seconds since the first call, measured with hipEvents
  void CheckErrors();

 private:
  Device() = default;
  bool initialized_ = false;
  hipDeviceProp_t prop_{};
  hipEvent_t start_ = nullptr;
};

// ------------------------------------------------------------------ boxes
class SimulationBox {
 public:
  virtual ~SimulationBox() = default;
  virtual float GetVolume() const = 0;
  virtual float GetLength(int dir) const = 0;
  virtual std::string GetInfoString(unsigned precision) const = 0;
  // RUMD's float[8] {cubic, Lx, Ly, Lz, 1/Lx, 1/Ly, 1/Lz, shift}
  const float* GetBox8() const { return box8_; }

 protected:
  float box8_[8] = {0, 0, 0, 0, 0, 0, 0, 0};
};

// RectangularSimulationBox (include/rumd/SimulationBox.h:96-190).
class RectangularSimulationBox : public SimulationBox {
 public:
  RectangularSimulationBox(float Lx, float Ly, float Lz);
  float GetVolume() const override;
  float GetLength(int dir) const override;
  std::string GetInfoString(unsigned precision) const override;
  void SetLengths(float Lx, float Ly, float Lz);
};

class Sample;
class ParticleData;

// ------------------------------------------------------------------ potentials
// Potential (include/rumd/Potential.h:16-63).
class Potential {
 public:
  virtual ~Potential() = default;
  virtual void CalcF(bool initialize, bool calc_stresses) = 0;
  virtual bool EnergyIncludedInParticleSum() const = 0;
  virtual void SetSample(Sample* s);
  virtual void Initialize() = 0;
  virtual void ResetInternalData() {}
  virtual double GetPotentialEnergy() { return 0.0; }
  virtual double GetVirial() { return 0.0; }
  // Bonds, angles: potentials whose terms come from the sample's topology.
  virtual bool IsMolecularPotential() const { return false; }
  void SetVerbose(bool vb) { verbose = vb; }
  void SetID_String(const std::string& s) { ID_String = s; }
  std::string GetID_String() const { return ID_String; }
  Sample* GetSample() const { return sample; }

 protected:
  Sample* sample = nullptr;
  bool verbose = true;
  std::string ID_String;
};

class PairPotential;

// NeighborList (include/rumd/NeighborList.h, src/NeighborList.cu).  Owns the
// list, its rebuild flags and the cell arrays of the sorted method.
class NeighborList {
 public:
  enum Method { DEFAULT, NONE, N2, SORT };
  NeighborList();
  void Initialize(Sample* s);
  void SetCutoffs(const PairPotential* pot);
  void SetSkin(float skin);
  float GetSkin() const { return skin_; }
  // Inner lists of the sorted method (no counterpart in the reference): the
  // force kernels read the list entries within rc + inner skin, pruned from
  // the list whenever a particle has moved 0.49 inner skin.  Zero (the
  // default): off; negative: skin/3.  Takes effect at the next allocation.
  void SetInnerSkin(float inner);
  float GetInnerSkin() const;
  bool HasInnerLists() const { return desc_.nbl16_in_dev != nullptr; }
  // Occupancy variant of the 12-6 tile force kernel: 0 (default, 5), 4, 5 or 6
  // waves per SIMD; the autotuner's tile mapping parameter.
  void SetTileWavesPerEU(unsigned w);
  unsigned GetTileWavesPerEU() const { return tile_waves_; }
  float GetMaxCutoff() const { return max_cutoff_; }
  void SetNB_Method(const std::string& method);
  std::string GetNB_Method() const;
  Method GetResolvedMethod() const;
  void SetMaxNumNbrs(unsigned m);
  void SetAllocateMax(bool b) { allocate_max_ = b; }
  unsigned GetActualMaxNumNbrs();
  int GetRebuildRequired();
  unsigned GetNumberOfBuilds();
  // Marks the list stale: the next UpdateNBlist rebuilds (ResetNeighborList).
  void ResetNeighborList();
  // Launches the device-side check and build chain (UpdateNBlist).
  void UpdateNBlist();
  // Allocates buffers now (before a graph capture, which may not allocate).
  void EnsureAllocated();
  // Called after an integrator step that ran this list's skin check fused.
  void NoteFusedCheck() { fused_check_pending_ = true; }
  void ResetFusedCheck() { fused_check_pending_ = false; }
  bool FusedCheckPending() const { return fused_check_pending_; }
  // True when the next sorted rebuild must be a full one (rumdx_nblist_build_sorted).
  bool NeedsFullSort() const { return GetResolvedMethod() == SORT && !UseTileResort(); }
  bool UseTileResort() const { return sort_valid_ && last_max_tile_ <= RUMDX_RESORT_TILE_MAX / 2; }
  void InvalidateSortOrder();
  // Every sorted rebuild sorts from scratch while set (for timing the two paths).
  void SetForceFullSort(bool on) { force_full_sort_ = on; }
  // Reads the flags; grows the list after an overflow.  Synchronises.
  bool HandleOverflow();
  // Asks the next UpdateNBlist to rebuild (device flag, on the sample's stream).
  void RequestRebuild();
  // Exclusions (NeighborList::SetExclusion, src/NeighborList.cu:791-850):
  // pairs of original particle ids that this list leaves out.  The lists are
  // rebuilt with them at the next update; method "none" then runs as N2.
  void SetExclusion(unsigned i, unsigned j);
  void AddExclusions(const std::vector<uint32_t>& pairs);
  void ClearExclusions();
  bool HasExclusions() const { return num_exclusions_ > 0; }
  unsigned GetNumberOfExclusions() const { return num_exclusions_; }
  bool IsSorted() const { return GetResolvedMethod() == SORT; }
  void DownloadCounts(std::vector<uint32_t>& out);
  // LDS halo capacity of the tile kernels, in particles (0 = estimate from
  // the density); tiles with larger halos compute from global memory.
  void SetHaloCapacity(unsigned cap);
  void SizeTileHalo();
  // Lanes sharing one particle's list in the force kernels: the sample's tp.
  unsigned GetLanesPerParticle() const { return desc_.lanes_per_particle; }
  unsigned GetHaloCapacity() const { return halo_cap_; }
  // Tiles of the last sorted build: {tiles, staged (halo in LDS), fast
  // (staged, pairs take the FP32 difference), wrapping (halo crosses the
  // box), shifted (wrapping and fast: staged in the shifted frame), largest
  // halo}.  Empty for the other methods.
  std::vector<uint32_t> TileStats();
  // The list in RUMD's format (u32 slots, column-major), for checks and tools.
  std::vector<uint32_t> ExportList(std::vector<uint32_t>& counts);
  rumdx_nblist* Descriptor() { return &desc_; }

 private:
  // Downloads the flags after gathering the last sorted build's statistics.
  void ReadFlags(uint32_t* f);
  void Allocate();
  void Settle();
  void Prune();
  Sample* sample_ = nullptr;
  Method method_ = DEFAULT;
  float skin_ = 0.5f;  // reference default (src/NeighborList.cu:50)
  float max_cutoff_ = 0.f;
  unsigned max_nbrs_ = 0;
  unsigned user_max_nbrs_ = 0;
  unsigned grown_max_nbrs_ = 0;  // capacity after an overflow
  bool allocate_max_ = false;
  bool fused_check_pending_ = false;
  bool allocated_ = false;
  unsigned ntypes_ = 0;
  std::vector<float> cutoffs_;  // rc_ij
  DeviceArray<uint32_t> nbl_, num_nbrs_, flags_, cell_rank_, cell_start_, cell_count_, slot_cell_;
  DeviceArray<uint32_t> tile_first_rank_, tile_origin_;
  DeviceArray<uint16_t> nbl16_;
  uint32_t halo_cap_ = 0;  // LDS halo capacity of the tile kernels (particles); 0 = estimate
  // Particles are in the cell order of this list's last sorted build (and the
  // home cell starts describe it), so a rebuild may re-sort tile by tile
  // (rumdx_nblist_rebuild_sorted); false until a full sorted build is queued.
  bool sort_valid_ = false;
  bool force_full_sort_ = false;
  uint32_t last_max_tile_ = 0;  // most particles in one tile, last synchronised build
  DeviceArray<float> last_r_, cut_skin2_;
  float inner_skin_ = 0.f;  // 0: off, < 0: skin/3
  unsigned tile_waves_ = 0;
  DeviceArray<uint16_t> nbl16_in_;
  DeviceArray<uint32_t> num_in_;
  DeviceArray<float> prune_r_, cut_in2_;
  DeviceArray<char> sort_scratch_;
  DeviceArray<uint16_t> lane_order_;
  DeviceArray<uint32_t> tile_halo_;  // ntiles x halo_cap
  DeviceArray<uint32_t> tile_info_;  // ntiles x 4
  rumdx_nblist desc_{};
  uint32_t allocated_nvp_ = 0;
  void UploadExclusions();
  std::vector<std::vector<uint32_t>> excl_;  // per original id
  unsigned num_exclusions_ = 0;
  DeviceArray<uint32_t> excl_start_, excl_ids_;
};

// PairPotential (include/rumd/PairPotential.h:33-128, src/PairPotential.cu).
class PairPotential : public Potential {
 public:
  enum CutoffMethod { NS = 0, SP = 1, SF = 2 };
  explicit PairPotential(CutoffMethod m);
  void CalcF(bool initialize, bool calc_stresses) override;
  bool EnergyIncludedInParticleSum() const override { return true; }
  void Initialize() override;
  void ResetInternalData() override { neighborList.ResetNeighborList(); }
  double GetPotentialEnergy() override;
  // Parameter rows: (Rcut, A12, A6, ...); shifts appended (CopyParamsToGPU).
  void SetParamsRow(unsigned i, unsigned j, const std::vector<float>& row);
  float GetPotentialParameter(unsigned i, unsigned j, unsigned idx) const;
  std::vector<float> GetPotentialParameters(unsigned num_types) const;
  // The full device table (12 floats per type pair, shifts included).
  std::vector<float> GetParamsTable(unsigned num_types) const;
  void SetNbListSkin(float skin) { neighborList.SetSkin(skin); }
  float GetNbListSkin() const { return neighborList.GetSkin(); }
  float GetMaxCutoff() const { return neighborList.GetMaxCutoff(); }
  void SetNB_Method(const std::string& m) { neighborList.SetNB_Method(m); }
  std::string GetNB_Method() const { return neighborList.GetNB_Method(); }
  void SetNbListAllocateMax(bool b) { neighborList.SetAllocateMax(b); }
  void SetNbMaxNumNbrs(unsigned m) { neighborList.SetMaxNumNbrs(m); }
  unsigned GetActualMaxNumNbrs() { return neighborList.GetActualMaxNumNbrs(); }
  int GetNbListRebuildRequired() { return neighborList.GetRebuildRequired(); }
  unsigned GetNumberOfNbListBuilds() { return neighborList.GetNumberOfBuilds(); }
  int GetCutoffMethod() const { return cutoff_method_; }
  void SetAssumeNewton3(bool b) { assume_newton3_ = b; }
  // Average duration (ms) of the force kernel alone over `reps` launches on
  // the sample's stream, bracketed by HIP events (the list is brought up to
  // date first and not rebuilt in between).
  float TimeForceKernel(unsigned reps);
  // Average duration (ms) of one full list rebuild (skin check, sort chain,
  // build) over `reps` forced rebuilds of the current positions, bracketed
  // by HIP events on the sample's stream.
  // Average duration (ms) of a forced rebuild: the tile re-sort path, or with
  // full_sort the full counting sort (each rep includes the skin check launch).
  float TimeNbListBuild(unsigned reps, bool full_sort = false);
  // All-FP64 pair arithmetic (rumdx_pair_lj_fp64): the precision study's
  // comparison path; FP32 pair arithmetic otherwise.
  void SetFP64Forces(bool on);
  bool GetFP64Forces() const { return fp64_forces_; }
  // True when a step of a sample whose only force term is this potential can
  // run as one kernel (rumdx_pair_lj_nvt): the 12-6 form in FP32 over an
  // allocated list of the sorted method whose tiles fit the sample's
  // workspace.
  bool CanFuseStep();
  // CalcF(true, false) and the NVT step in one kernel; the new positions go to
  // the sample's second position array, which then becomes the current one.
  void CalcFIntegrateNVT(rumdx_nh_state* nh_dev, float dt, unsigned thermostat_on_mask, float omega2, unsigned dof);
  // Average duration (ms) of the fused step kernel alone over `reps`
  // launches, bracketed by HIP events like TimeForceKernel.  The launches
  // advance copies of the velocities, images and low position parts, write
  // their positions to a scratch array and raise their rebuild requests in a
  // scratch copy of the list flags, so the sample's trajectory is untouched;
  // the thermostat is off in them, which leaves out only the per-workgroup
  // kinetic-energy sum and the one-workgroup thermostat update after it.
  float TimeFusedStepKernel(unsigned reps, float dt);
  // Mean and maximum current list length (synchronises).
  std::vector<double> GetNeighborListStats();
  NeighborList& GetNeighborList() { return neighborList; }
  // Shift constants u(rc), s(rc) rc of one parameter row.
  virtual void ComputeShifts(std::vector<float>& row) const = 0;
  // 0 for the 12-6 family (rumdx_pair_lj), else the RUMDX_POT_* form served
  // by rumdx_pair_form.
  virtual int GetPotentialForm() const { return 0; }
  unsigned NumTypesInTable() const { return table_types_; }

 protected:
  void UploadParams();
  NeighborList neighborList;
  CutoffMethod cutoff_method_;
  bool assume_newton3_ = true;
  bool fp64_forces_ = false;
  void LaunchForce(const rumdx_particles* p, bool initialize, bool calc_stresses, hipStream_t st, const char* method);
  std::map<std::pair<unsigned, unsigned>, std::vector<float>> params_map_;
  DeviceArray<float> d_params_;
  unsigned table_types_ = 0;
  DeviceArray<float> pe_f_, pe_w_;  // private arrays for GetPotentialEnergy
};

// Pot_IPL_12_6 (include/rumd/PairPotential.h:134-169).
class Pot_IPL_12_6 : public PairPotential {
 public:
  explicit Pot_IPL_12_6(CutoffMethod m) : PairPotential(m) { SetID_String("potIPL_12_6"); }
  void SetParams(unsigned i, unsigned j, float A12, float A6, float Rcut);
  void ComputeShifts(std::vector<float>& row) const override;
};

// Pot_LJ_12_6 (include/rumd/PairPotential.h:184-208).
class Pot_LJ_12_6 : public Pot_IPL_12_6 {
 public:
  explicit Pot_LJ_12_6(CutoffMethod m) : Pot_IPL_12_6(m) { SetID_String("potLJ_12_6"); }
  void SetParams(unsigned i, unsigned j, float Sigma, float Epsilon, float Rcut);
};

// Pot_gLJ_m12_n6 (include/rumd/PairPotential.h:312-330): 12-6 LJ with its
// minimum, -Epsilon, at Sigma.
class Pot_gLJ_m12_n6 : public Pot_IPL_12_6 {
 public:
  explicit Pot_gLJ_m12_n6(CutoffMethod m) : Pot_IPL_12_6(m) { SetID_String("potLJA"); }
  void SetParams(unsigned i, unsigned j, float Sigma, float Epsilon, float Rcut);
};

// Pair potentials of the other forms (rumdx_pair_form, one form per class):
// parameter rows as the reference's SetParams stores them, cut-off shifts from
// the form at rc.
class FormPairPotential : public PairPotential {
 public:
  FormPairPotential(CutoffMethod m, int form) : PairPotential(m), form_(form) {}
  int GetPotentialForm() const override { return form_; }
  void ComputeShifts(std::vector<float>& row) const override;

 protected:
  void Report(const char* what) const;
  int form_;
};

// Pot_LJ_12_6_smooth_2_4 (include/rumd/PairPotential.h:218-256).
class Pot_LJ_12_6_smooth_2_4 : public FormPairPotential {
 public:
  explicit Pot_LJ_12_6_smooth_2_4(CutoffMethod m);
  void SetParams(unsigned i, unsigned j, float Sigma, float Epsilon, float Rcut);
};

// Pot_gLJ_m_n (include/rumd/PairPotential.h:264-304).
class Pot_gLJ_m_n : public FormPairPotential {
 public:
  Pot_gLJ_m_n(float m, float n, CutoffMethod cm);
  void SetParams(unsigned i, unsigned j, float Sigma, float Epsilon, float Rcut);
  float m, n;
};

// Pot_Gauss (include/rumd/PairPotential.h:383-414).
class Pot_Gauss : public FormPairPotential {
 public:
  explicit Pot_Gauss(CutoffMethod m);
  void SetParams(unsigned i, unsigned j, float Sigma, float Epsilon, float Rcut);
};

// Pot_Buckingham (include/rumd/PairPotential.h:422-461).
class Pot_Buckingham : public FormPairPotential {
 public:
  explicit Pot_Buckingham(CutoffMethod m);
  void SetParams(unsigned i, unsigned j, float rm, float Alpha, float Epsilon, float Rcut);
};

// Pot_IPL_n (include/rumd/PairPotential.h:600-636).
class Pot_IPL_n : public FormPairPotential {
 public:
  Pot_IPL_n(float n, CutoffMethod cm);
  void SetParams(unsigned i, unsigned j, float Sigma, float Epsilon, float Rcut);
  float n;
};

// Pot_Yukawa (include/rumd/PairPotential.h:795-840).
class Pot_Yukawa : public FormPairPotential {
 public:
  explicit Pot_Yukawa(CutoffMethod m);
  void SetParams(unsigned i, unsigned j, float Sigma, float Epsilon, float Rcut);
};

// Pot_Exp (include/rumd/PairPotential.h:1019-1052).
class Pot_Exp : public FormPairPotential {
 public:
  explicit Pot_Exp(CutoffMethod m);
  void SetParams(unsigned i, unsigned j, float Sigma, float Epsilon, float Rcut);
};

// Pot_LJ_12_6_Gauss (include/rumd/PairPotential.h:336-378).
class Pot_LJ_12_6_Gauss : public FormPairPotential {
 public:
  explicit Pot_LJ_12_6_Gauss(CutoffMethod m);
  void SetParams(unsigned i, unsigned j, float Sigma, float Epsilon, float Rcut, float Sigma0, float Epsilon0,
                 float r0);
};

// Pot_IPL_12, Pot_IPL_9, Pot_IPL_18 (include/rumd/PairPotential.h:469-593):
// one class per fixed exponent.
class Pot_IPL_fixed : public FormPairPotential {
 public:
  Pot_IPL_fixed(CutoffMethod m, int form, float n, const char* id);
  void SetParams(unsigned i, unsigned j, float Sigma, float Epsilon, float Rcut);

 private:
  float n_;
};
class Pot_IPL_12 : public Pot_IPL_fixed {
 public:
  explicit Pot_IPL_12(CutoffMethod m) : Pot_IPL_fixed(m, RUMDX_POT_IPL_12, 12.f, "potIPL12") {}
};
class Pot_IPL_9 : public Pot_IPL_fixed {
 public:
  explicit Pot_IPL_9(CutoffMethod m) : Pot_IPL_fixed(m, RUMDX_POT_IPL_9, 9.f, "potIPL9") {}
};
class Pot_IPL_18 : public Pot_IPL_fixed {
 public:
  explicit Pot_IPL_18(CutoffMethod m) : Pot_IPL_fixed(m, RUMDX_POT_IPL_18, 18.f, "potIPL18") {}
};

// Pot_IPL6_LN (include/rumd/PairPotential.h:644-680).
class Pot_IPL6_LN : public FormPairPotential {
 public:
  explicit Pot_IPL6_LN(CutoffMethod m);
  void SetParams(unsigned i, unsigned j, float Sigma, float Epsilon, float Rcut);
};

// Pot_LJ_LN (include/rumd/PairPotential.h:689-733).
class Pot_LJ_LN : public FormPairPotential {
 public:
  explicit Pot_LJ_LN(CutoffMethod m);
  void SetParams(unsigned i, unsigned j, float Rm, float R0, float Epsilon, float Rcut);
};

// Pot_Girifalco (include/rumd/PairPotential.h:743-786).
class Pot_Girifalco : public FormPairPotential {
 public:
  explicit Pot_Girifalco(CutoffMethod m);
  void SetParams(unsigned i, unsigned j, float alpha, float beta, float R0, float Rcut);
};

// Pot_Dzugutov (include/rumd/PairPotential.h:846-900): no cut-off shift.
class Pot_Dzugutov : public FormPairPotential {
 public:
  Pot_Dzugutov();
  void SetParams(unsigned i, unsigned j, float a, float b, float c, float d, float A, float B);
};

// Pot_Fermi_Jagla (include/rumd/PairPotential.h:910-961).
class Pot_Fermi_Jagla : public FormPairPotential {
 public:
  explicit Pot_Fermi_Jagla(CutoffMethod m);
  void SetParams(unsigned i, unsigned j, float a, float epsilon0, float A0, float A1, float A2, float B0, float B1,
                 float B2, float n, float Rcut);
};

// Pot_Repulsive_Shoulder (include/rumd/PairPotential.h:969-1012).
class Pot_Repulsive_Shoulder : public FormPairPotential {
 public:
  explicit Pot_Repulsive_Shoulder(CutoffMethod m);
  void SetParams(unsigned i, unsigned j, float Sigma, float Epsilon, float k0, float Sigma1, float Rcut);
};

// Pot_SAAP (include/rumd/PairPotential.h:1060-1135).
class Pot_SAAP : public FormPairPotential {
 public:
  explicit Pot_SAAP(CutoffMethod m);
  void SetParams(unsigned i, unsigned j, float Sigma, float Epsilon, float a0, float a1, float a2, float a3, float a4,
                 float a5, float a6, float Rcut);
};

// Pot_ShiftedForceCoulomb (include/rumd/PairPotential.h:1143-1186): the
// shifted force is part of the form, so the class uses no cut-off shift.
class Pot_ShiftedForceCoulomb : public FormPairPotential {
 public:
  Pot_ShiftedForceCoulomb();
  void SetParams(unsigned i, unsigned j, float Epsilon, float Rcut);
};

// ------------------------------------------------------------------ molecules
// MoleculeData (include/rumd/MoleculeData.h, src/MoleculeData.cu:479-700): the
// bond and angle topology, held on the device as per-particle lists (CSR over
// original particle ids) for the gather kernels of rumdx_bond_harmonic and
// rumdx_angle_cossq, and the slot index they read particles through.
class MoleculeData {
 public:
  explicit MoleculeData(Sample* s) : sample_(s) {}
  // Reads the [ bonds ] (mol a b type) and [ angles ] (mol a b c type)
  // sections of a RUMD topology file; other sections are ignored.
  void ReadTopology(const std::string& filename);
  // Replaces the bonds: flat (a, b, type) triples of original particle ids.
  void SetBonds(const std::vector<uint32_t>& abt, const std::vector<uint32_t>& molecule = {});
  // Replaces the angles: flat (a, b, c, type) quadruples.
  void SetAngles(const std::vector<uint32_t>& abct);
  // Replaces the dihedrals: flat (a, b, c, d, type) quintuples.
  void SetDihedrals(const std::vector<uint32_t>& abcdt);
  unsigned GetNumberOfDihedrals() const { return static_cast<unsigned>(dihedrals_.size() / 5); }
  unsigned GetNumberOfDihedralTypes() const;
  const std::vector<uint32_t>& GetDihedrals() const { return dihedrals_; }
  const uint32_t* DihedralStart() const { return dihedral_start_.get(); }
  const uint32_t* DihedralAtoms() const { return dihedral_atoms_.get(); }
  unsigned GetNumberOfBonds() const { return static_cast<unsigned>(bonds_.size() / 3); }
  unsigned GetNumberOfAngles() const { return static_cast<unsigned>(angles_.size() / 4); }
  unsigned GetNumberOfMolecules() const { return num_mol_; }
  unsigned GetNumberOfBondTypes() const;
  unsigned GetNumberOfAngleTypes() const;
  const std::vector<uint32_t>& GetBonds() const { return bonds_; }
  const std::vector<uint32_t>& GetAngles() const { return angles_; }
  // Refreshes slot_of from the particles' unsorted index (on the sample's stream).
  const uint32_t* SlotIndex();
  const uint32_t* BondStart() const { return bond_start_.get(); }
  const uint32_t* BondOther() const { return bond_other_.get(); }
  const uint32_t* BondType() const { return bond_type_.get(); }
  const uint32_t* AngleStart() const { return angle_start_.get(); }
  const uint32_t* AngleAtoms() const { return angle_atoms_.get(); }

 private:
  void UploadBonds();
  void UploadAngles();
  void UploadDihedrals();
  Sample* sample_;
  unsigned num_mol_ = 0;
  std::vector<uint32_t> bonds_, angles_, dihedrals_;
  uint64_t slot_epoch_ = 0;  // sample epoch of the last slot index
  DeviceArray<uint32_t> bond_start_, bond_other_, bond_type_, angle_start_, angle_atoms_, slot_of_;
  DeviceArray<uint32_t> dihedral_start_, dihedral_atoms_;
};

// BondHarmonic (include/rumd/BondPotential.h, src/BondPotential.cu:70-135):
// u = k/2 (r - l0)^2 per bond of the sample's topology whose type has
// parameters here (a bond type belongs to the bond potential whose SetParams
// named it, as in the reference).
class BondHarmonic : public Potential {
 public:
  BondHarmonic() { SetID_String("bondHarmonic"); }
  void SetParams(unsigned bond_type, float bond_length, float stiffness, bool exclude);
  void CalcF(bool initialize, bool calc_stresses) override;
  bool EnergyIncludedInParticleSum() const override { return true; }
  void Initialize() override {}
  double GetPotentialEnergy() override;
  bool IsMolecularPotential() const override { return true; }
  std::vector<float> GetParams() const { return params_; }
  // Bond types whose pairs the non-bonded potentials skip.
  std::vector<unsigned> GetExcludedTypes() const;
  // BondPotential::SetExclusions (src/BondPotential.cu:54-66): the pairs of
  // bonds of excluded types leave the pair potential's neighbour list.
  void SetExclusions(PairPotential* non_bond_pot);

 private:
  void Launch(const rumdx_particles* p, bool initialize, bool calc_stresses, const char* method);

 protected:
  bool fene_ = false;
  std::vector<float> params_;  // (l0 or R0, k) per bond type; k = 0: not this potential's type
  std::vector<bool> exclude_;
  DeviceArray<float> d_params_;
  DeviceArray<float> pe_f_, pe_w_;
};

// BondFENE (include/rumd/BondPotential.h:47-56, src/BondPotential.cu):
// u = -k R0^2/2 ln(1 - r^2/R0^2); SetParams(bond_type, R0, k, exclude).
class BondFENE : public BondHarmonic {
 public:
  BondFENE() {
    SetID_String("bondFene");
    fene_ = true;
  }
};

// DihedralRyckaert (include/rumd/DihedralPotential.h, src/DihedralPotential.cu:78-110):
// u = sum_k c_k cos^k(psi) per dihedral of the topology.
class DihedralRyckaert : public Potential {
 public:
  DihedralRyckaert() { SetID_String("DihedralRyckaert"); }
  void SetParams(unsigned dihedral_type, const std::vector<float>& coeffs);
  void CalcF(bool initialize, bool calc_stresses) override;
  bool EnergyIncludedInParticleSum() const override { return true; }
  void Initialize() override {}
  double GetPotentialEnergy() override;
  bool IsMolecularPotential() const override { return true; }
  // DihedralPotential::SetExclude / SetExclusions: the end atoms (a, d) leave
  // the pair potential's list (on by default).
  void SetExclude(bool exclude) { exclude_dihedral_ = exclude; }
  void SetExclusions(PairPotential* non_bond_pot);

 protected:
  size_t num_params_ = 6;  // coefficients SetParams takes (DihedralPotential's num_dihedral_params)
  bool periodic_ = false;  // PeriodicDihedral

 private:
  void Launch(const rumdx_particles* p, bool initialize, bool calc_stresses, const char* method);
  std::vector<float> params_;  // c_0..c_5 per type
  bool exclude_dihedral_ = true;
  DeviceArray<float> d_params_;
  DeviceArray<float> pe_f_, pe_w_;
};

// PeriodicDihedral (src/DihedralPotential.cu:265-458): params (psi0, k1, k2,
// k3), u = 1/2 sum_m k_m (1 - cos m (psi - psi0)) with the reference's force,
// which is -sin psi times the gradient force (a property of that kernel,
// reproduced here).
class PeriodicDihedral : public DihedralRyckaert {
 public:
  PeriodicDihedral() {
    SetID_String("PeriodicDihedral");
    num_params_ = 4;
    periodic_ = true;
  }
};

// AngleCosSq (include/rumd/AnglePotential.h, src/AnglePotential.cu:69-98):
// u = k/2 (cos theta - cos theta0)^2 per angle.
class AngleCosSq : public Potential {
 public:
  AngleCosSq() { SetID_String("AngleCosSq"); }
  void SetParams(unsigned angle_type, float theta0, float ktheta);
  void CalcF(bool initialize, bool calc_stresses) override;
  bool EnergyIncludedInParticleSum() const override { return true; }
  void Initialize() override {}
  double GetPotentialEnergy() override;
  bool IsMolecularPotential() const override { return true; }
  std::vector<float> GetParams() const { return params_; }
  // AnglePotential::SetExclude / SetExclusions (src/AnglePotential.cu:63-66):
  // the end atoms (a, c) of every angle leave the pair potential's list
  // (on by default, as in the reference).
  void SetExclude(bool exclude) { exclude_angle_ = exclude; }
  void SetExclusions(PairPotential* non_bond_pot);

 protected:
  bool squared_ = false;  // AngleSq: u = k/2 (theta - theta0)^2

 private:
  void Launch(const rumdx_particles* p, bool initialize, bool calc_stresses, const char* method);
  std::vector<float> params_;  // (theta0, k) per angle type
  bool exclude_angle_ = true;
  DeviceArray<float> d_params_;
  DeviceArray<float> pe_f_, pe_w_;
};

// AngleSq (src/AnglePotential.cu:216-373): u = k/2 (theta - theta0)^2 per
// angle, with AngleCosSq's parameters, exclusions and energy placement.
class AngleSq : public AngleCosSq {
 public:
  AngleSq() {
    SetID_String("AngleSq");
    squared_ = true;
  }
};

// ------------------------------------------------------------------ integrators
// Integrator (include/rumd/Integrator.h:18-70).
class Integrator {
 public:
  virtual ~Integrator() = default;
  virtual void SetSample(Sample* s) { S = s; }
  virtual void Integrate() = 0;
  virtual std::string GetInfoString(unsigned precision) const = 0;
  virtual void InitializeFromInfoString(const std::string& info, bool verbose = true) = 0;
  virtual void SetMomentumToZero() {}
  virtual double GetKineticEnergy(bool copy = true) = 0;
  virtual void SetTimeStep(float dt) { timeStep = dt; }
  virtual float GetTimeStep() const { return timeStep; }
  std::vector<double> GetTotalMomentum();

 protected:
  Sample* S = nullptr;
  float timeStep = 1.f;
};

// IntegratorNVT (include/rumd/IntegratorNVT.h, src/IntegratorNVT.cu): NVE
// leapfrog when built with only a time step, Nose-Hoover otherwise.
class IntegratorNVT : public Integrator {
 public:
  explicit IntegratorNVT(float timeStep);  // NVE
  IntegratorNVT(float timeStep, double targetTemperature, double thermostatRelaxationTime);
  void SetSample(Sample* s) override;
  void Integrate() override;
  std::string GetInfoString(unsigned precision) const override;
  void InitializeFromInfoString(const std::string& info, bool verbose = true) override;
  void SetMomentumToZero() override;
  double GetKineticEnergy(bool copy = true) override;
  void SetTimeStep(float dt) override;
  void SetRelaxationTime(float tau);
  float GetRelaxationTime() const { return relaxation_time_; }
  void SetTargetTemperature(double T, bool reset = true);
  double GetTargetTemperature();
  void SetThermostatState(double ps);
  double GetThermostatState();
  double GetThermostatKineticEnergy();  // last FP64 sum of m v^2 / 2
  void SetThermostatOn(bool on);
  void SetThermostatOnParticleType(unsigned type, bool on);
  unsigned GetThermostatOnMask() const { return thermostat_on_; }
  void SetTemperatureChangePerTimeStep(double dT);
  // The step with the forces of `pp` computed in the same kernel
  // (PairPotential::CalcFIntegrateNVT): CalcF and Integrate of Simulation.Run.
  void IntegrateFused(PairPotential& pp);

 private:
  void PushState();
  void PullState();
  unsigned thermostat_on_ = 0;
  float relaxation_time_ = 0.2f;
  rumdx_nh_state host_state_{0.0, 0.0, 0.0, 0.0};
  DeviceArray<rumdx_nh_state> d_state_;
};

// ------------------------------------------------------------------ particle data
// ParticleData (include/rumd/ParticleData.h): device arrays in slot order,
// host mirrors in original particle order.
class ParticleData {
 public:
  void SetNumberOfParticles(unsigned n, unsigned pb, bool mixed_positions);
  unsigned GetNumberOfParticles() const { return n_; }
  unsigned GetNumberOfVirtualParticles() const { return nvp_; }
  unsigned GetNumberOfTypes() const { return static_cast<unsigned>(masses_.size()); }
  unsigned GetNumberThisType(unsigned t) const;
  void SetNumberOfTypes(unsigned nt);
  void SetMass(unsigned type, float mass);
  float GetMass(unsigned type) const;
  bool MixedPositions() const { return mixed_; }
  // host <-> device in original order (the unsorted index is applied here)
  void SetPositions(const float* xyz, hipStream_t st);
  // FP64 positions: the float nearest each coordinate, and in mixed mode the
  // remainder as the low-order part (so GetPositionsFP64 returns them)
  void SetPositionsFP64(const double* xyz, hipStream_t st);
  void SetTypes(const int32_t* types, hipStream_t st);
  void SetVelocities(const float* xyz, hipStream_t st);
  void SetImages(const float* xyz, hipStream_t st);
  void SetForces(const float* f4, hipStream_t st);
  void Get(const char* what, float* out4, hipStream_t st);  // "r","v","f","w","im","sts": n x 4, original order
  void GetTypes(int32_t* out, hipStream_t st);
  void ResetSorting(hipStream_t st);
  rumdx_particles* Descriptor() { return &desc_; }
  std::vector<uint32_t> UnsortedIndex(hipStream_t st);
  // The second position array of fused steps (rumdx_pair_lj_nvt writes the
  // new positions there): allocated on demand, outside graph capture.
  void EnsureNextPositions();
  bool HasNextPositions() const { return d_r_next.size() == d_r.size() && d_r.size() != 0; }
  float* NextPositions() const { return d_r_next.get(); }
  // Makes the second array the current one (d_r and the descriptor follow).
  void SwapPositions();
  // Brings the positions back into the array `home` (one of the two) by a
  // device copy on `st` when they are in the other one.
  void MovePositionsTo(float* home, hipStream_t st);

  DeviceArray<float> d_r, d_v, d_f, d_w, d_im, d_sts, d_rlo, d_r_next;
  DeviceArray<uint32_t> d_unsorted_index;

 private:
  void UpdateMassesOnDevice(hipStream_t st);
  void Scatter(const std::vector<float>& slot4, float* out4, hipStream_t st);
  std::vector<float> GatherToSlots(const float* orig4, hipStream_t st);
  unsigned n_ = 0, nvp_ = 0;
  bool mixed_ = false;
  std::vector<float> masses_;
  std::vector<int32_t> types_;  // original order
  rumdx_particles desc_{};
};

// ------------------------------------------------------------------ Sample
// Sample (include/rumd/Sample.h, src/Sample.cu): particle data, box, potential
// list and integrator; CalcF and the FP64 sums.  One HIP stream per sample.
class Sample {
 public:
  Sample(unsigned pb, unsigned tp);
  ~Sample();
  Sample(const Sample&) = delete;
  Sample& operator=(const Sample&) = delete;

  void SetNumberOfParticles(unsigned n);
  void SetSimulationBox(SimulationBox* box);
  SimulationBox* GetSimulationBox() const { return box_; }
  void SetPotential(Potential* p);
  void AddPotential(Potential* p);
  void SetIntegrator(Integrator* itg);
  Integrator* GetIntegrator() const { return itg_; }
  void SetVerbose(bool vb) { verbose_ = vb; }
  bool GetVerbose() const { return verbose_; }
  void SetPB_TP(unsigned pb, unsigned tp);
  unsigned GetParticlesPerBlock() const { return pb_; }
  unsigned GetThreadsPerParticle() const { return tp_; }
  void SetMixedPrecisionPositions(bool on);  // FP64 position bookkeeping
  bool GetMixedPrecisionPositions() const { return mixed_; }

  void CalcF(bool calc_stresses = false);
  double GetPotentialEnergy();
  double GetVirial();
  std::vector<double> GetStress();
  unsigned GetNumberOfDOFs() const;
  void SetNumberOfDOFs(unsigned dofs) { user_dofs_ = dofs; }
  unsigned GetNumberOfParticles() const { return pd_.GetNumberOfParticles(); }
  unsigned GetNumberOfVirtualParticles() const { return pd_.GetNumberOfVirtualParticles(); }
  unsigned GetNumberOfTypes() const { return pd_.GetNumberOfTypes(); }
  unsigned GetNumberThisType(unsigned t) const { return pd_.GetNumberThisType(t); }
  void SetMass(unsigned type, float mass);
  float GetMass(unsigned type) const { return pd_.GetMass(type); }
  float GetMeanMass() const;
  void SortParticles();

  // Runs nsteps of [CalcF, Integrate] with blocks of `block` steps captured as
  // HIP graphs; momentum is reset before every step whose index is a
  // multiple of momentum_interval (0 = never), as Simulation.Run does.
  // Blocks start at first_step and end early at a reset step, so a reset is
  // always the first step of a block (see RunSegments).
  void RunSteps(unsigned long first_step, unsigned nsteps, unsigned block, unsigned momentum_interval);
  // The blocks RunSteps would run for these arguments: (steps, reset at the
  // first step) each; blocks of one step run without a graph.
  std::vector<std::pair<unsigned, bool>> RunSegments(unsigned long first_step, unsigned nsteps, unsigned block,
                                                     unsigned momentum_interval) const;
  // Captures and instantiates every graph RunSteps needs for these arguments
  // without running any step, so that a later RunSteps launches graphs only
  // (graphs stay valid until InvalidateGraphs).  Returns the number captured.
  unsigned PrepareRunSteps(unsigned long first_step, unsigned nsteps, unsigned block, unsigned momentum_interval);
  void InvalidateGraphs();
  // Steps of a sample whose only force term is a 12-6 pair potential over a
  // sorted list, integrated by IntegratorNVT, run the force and the
  // integrator in one kernel (on by default; no counterpart in the
  // reference, whose steps always launch them apart).
  void SetFusedSteps(bool on);
  bool GetFusedSteps() const { return fused_steps_; }
  // The potential whose kernel takes the whole step, or null.
  PairPotential* FusedStepPotential();
  // PairPotential::TimeFusedStepKernel with the integrator's time step.
  float TimeFusedStepKernel(unsigned reps);

  ParticleData& GetParticleData() { return pd_; }
  hipStream_t GetStream() const { return stream_; }
  void* Workspace() { return workspace_.get(); }
  double* ScalarOut() { return scalars_.get(); }
  const std::vector<Potential*>& GetPotentials() const { return potentials_; }
  // Counts events that may change which slot holds which particle (force
  // calls, sorted list updates); a slot index computed at the same count is
  // still valid.
  uint64_t GetSlotEpoch() const { return slot_epoch_; }
  void BumpSlotEpoch() { ++slot_epoch_; }
  // Topology of the sample's molecules (Sample::ReadMoleculeData); null until read.
  void ReadMoleculeData(const std::string& topology_file);
  MoleculeData* GetMoleculeData();
  MoleculeData* CreateMoleculeData();
  NeighborList* FusedCheckList();
  void Synchronize();

 private:
  // A captured step block.  It reads the positions from the array that was
  // current when it was recorded (fused steps alternate two arrays) and ends
  // in the same one.
  struct Block {
    hipGraphExec_t exec;
    float* r_home;
  };
  void PrepareRun();
  void StepOnce(bool zero_momentum);
  Block CaptureBlock(unsigned nsteps, bool reset_first);
  unsigned pb_ = 0, tp_ = 0;
  unsigned user_dofs_ = 0;
  bool verbose_ = true;
  bool mixed_ = false;
  SimulationBox* box_ = nullptr;
  std::vector<Potential*> potentials_;
  Integrator* itg_ = nullptr;
  ParticleData pd_;
  hipStream_t stream_ = nullptr;
  DeviceArray<char> workspace_;
  DeviceArray<double> scalars_;
  // captured step blocks: key = (block size, momentum reset at first step)
  std::map<std::pair<unsigned, bool>, Block> graphs_;
  bool fused_steps_ = true;
  std::unique_ptr<MoleculeData> molecules_;
  uint64_t slot_epoch_ = 1;
};
