// MoleculeData and the intramolecular potentials BondHarmonic, AngleCosSq.
#include <cmath>
#include <cstdio>
#include <fstream>
#include <sstream>

#include "rumd_host.h"

namespace {

// CSR over original ids from per-particle entry lists.
void build_csr(unsigned n, const std::vector<std::vector<uint32_t>>& per, unsigned width, std::vector<uint32_t>& start,
               std::vector<uint32_t>& flat) {
  start.assign(n + 1, 0u);
  flat.clear();
  for (unsigned o = 0; o < n; ++o) {
    start[o] = static_cast<uint32_t>(flat.size() / width);
    flat.insert(flat.end(), per[o].begin(), per[o].end());
  }
  start[n] = static_cast<uint32_t>(flat.size() / width);
}

template <class T>
void upload_vec(DeviceArray<T>& d, const std::vector<T>& h, hipStream_t st) {
  d.resize(std::max<size_t>(h.size(), 1));
  if (!h.empty()) d.upload(h.data(), h.size(), st);
}

// Sums f.w of a private force array (the potential's own energy).
double sum_energy(Sample* s, const float* f, const char* cls) {
  const hipStream_t st = s->GetStream();
  rumdx_check(rumdx_reduce_f4(f, s->GetNumberOfParticles(), 8u, s->ScalarOut(), s->Workspace(), st), cls,
              "GetPotentialEnergy");
  double out[4];
  hip_check(hipMemcpyAsync(out, s->ScalarOut(), sizeof(out), hipMemcpyDeviceToHost, st), cls, "GetPotentialEnergy");
  hip_check(hipStreamSynchronize(st), cls, "GetPotentialEnergy");
  return out[3];
}

}  // namespace

// ------------------------------------------------------------------ MoleculeData
// MoleculeData::ReadTopology / GetBondsFromTopology / GetAnglesFromTopology
// (src/MoleculeData.cu:479-660): after a "[ bonds ]" line and one comment line,
// rows "mol a b type" until the next section; "[ angles ]" rows "mol a b c type".
void MoleculeData::ReadTopology(const std::string& filename) {
  std::ifstream in(filename);
  if (!in) throw RUMD_Error("MoleculeData", "ReadTopology", "Couldn't open topology file " + filename);
  std::vector<uint32_t> bonds, angles, dihedrals, mols;
  std::string line, section;
  while (std::getline(in, line)) {
    const size_t b = line.find_first_not_of(" \t\r");
    if (b == std::string::npos) continue;
    if (line[b] == '[') {
      const size_t e = line.find(']', b);
      std::istringstream t(line.substr(b + 1, e == std::string::npos ? std::string::npos : e - b - 1));
      t >> section;
      continue;
    }
    if (line[b] == ';' || line[b] == '#') continue;
    std::istringstream row(line);
    if (section == "bonds") {
      unsigned m, a, c, t;
      if (!(row >> m >> a >> c >> t)) throw RUMD_Error("MoleculeData", "ReadTopology", "Format not correct: " + line);
      bonds.insert(bonds.end(), {a, c, t});
      mols.push_back(m);
    } else if (section == "angles") {
      unsigned m, a, c, d, t;
      if (!(row >> m >> a >> c >> d >> t))
        throw RUMD_Error("MoleculeData", "ReadTopology", "Format not correct: " + line);
      angles.insert(angles.end(), {a, c, d, t});
    } else if (section == "dihedrals") {
      unsigned m, a, c, d, e, t;
      if (!(row >> m >> a >> c >> d >> e >> t))
        throw RUMD_Error("MoleculeData", "ReadTopology", "Format not correct: " + line);
      dihedrals.insert(dihedrals.end(), {a, c, d, e, t});
    }
  }
  SetBonds(bonds, mols);
  SetAngles(angles);
  SetDihedrals(dihedrals);
  if (sample_->GetVerbose())
    printf("Topology file %s: %u molecules, %u bonds, %u angles, %u dihedrals.\n", filename.c_str(), num_mol_,
           GetNumberOfBonds(), GetNumberOfAngles(), GetNumberOfDihedrals());
}

void MoleculeData::SetBonds(const std::vector<uint32_t>& abt, const std::vector<uint32_t>& molecule) {
  if (abt.size() % 3) throw RUMD_Error("MoleculeData", "SetBonds", "bonds must be (a, b, type) triples");
  const unsigned n = sample_->GetNumberOfParticles();
  for (size_t k = 0; k < abt.size(); k += 3)
    if (abt[k] >= n || abt[k + 1] >= n || abt[k] == abt[k + 1])
      throw RUMD_Error("MoleculeData", "SetBonds", "bond refers to a particle outside the sample or to itself");
  bonds_ = abt;
  num_mol_ = 0;
  for (uint32_t m : molecule) num_mol_ = std::max(num_mol_, m + 1);
  UploadBonds();
}

void MoleculeData::SetAngles(const std::vector<uint32_t>& abct) {
  if (abct.size() % 4) throw RUMD_Error("MoleculeData", "SetAngles", "angles must be (a, b, c, type) quadruples");
  const unsigned n = sample_->GetNumberOfParticles();
  for (size_t k = 0; k < abct.size(); k += 4)
    if (abct[k] >= n || abct[k + 1] >= n || abct[k + 2] >= n || abct[k] == abct[k + 1] ||
        abct[k + 1] == abct[k + 2] || abct[k] == abct[k + 2])
      throw RUMD_Error("MoleculeData", "SetAngles", "angle must name three distinct particles of the sample");
  angles_ = abct;
  UploadAngles();
}

void MoleculeData::SetDihedrals(const std::vector<uint32_t>& abcdt) {
  if (abcdt.size() % 5) throw RUMD_Error("MoleculeData", "SetDihedrals", "dihedrals must be (a, b, c, d, type)");
  const unsigned n = sample_->GetNumberOfParticles();
  for (size_t k = 0; k < abcdt.size(); k += 5) {
    for (int u = 0; u < 4; ++u)
      for (int v = 0; v < 4; ++v)
        if (abcdt[k + u] >= n || (u < v && abcdt[k + u] == abcdt[k + v]))
          throw RUMD_Error("MoleculeData", "SetDihedrals", "dihedral must name four distinct particles of the sample");
  }
  dihedrals_ = abcdt;
  UploadDihedrals();
}

unsigned MoleculeData::GetNumberOfDihedralTypes() const {
  unsigned t = 0;
  for (size_t k = 4; k < dihedrals_.size(); k += 5) t = std::max(t, dihedrals_[k] + 1);
  return t;
}

void MoleculeData::UploadDihedrals() {
  const unsigned n = sample_->GetNumberOfParticles();
  std::vector<std::vector<uint32_t>> per(n);
  for (size_t k = 0; k < dihedrals_.size(); k += 5)
    for (int u = 0; u < 4; ++u)
      per[dihedrals_[k + u]].insert(per[dihedrals_[k + u]].end(), dihedrals_.begin() + k, dihedrals_.begin() + k + 5);
  std::vector<uint32_t> start, flat;
  build_csr(n, per, 5, start, flat);
  const hipStream_t st = sample_->GetStream();
  upload_vec(dihedral_start_, start, st);
  upload_vec(dihedral_atoms_, flat, st);
  if (slot_of_.size() != n) slot_of_.resize(n);
  slot_epoch_ = 0;
  sample_->InvalidateGraphs();
}

unsigned MoleculeData::GetNumberOfBondTypes() const {
  unsigned t = 0;
  for (size_t k = 2; k < bonds_.size(); k += 3) t = std::max(t, bonds_[k] + 1);
  return t;
}

unsigned MoleculeData::GetNumberOfAngleTypes() const {
  unsigned t = 0;
  for (size_t k = 3; k < angles_.size(); k += 4) t = std::max(t, angles_[k] + 1);
  return t;
}

void MoleculeData::UploadBonds() {
  const unsigned n = sample_->GetNumberOfParticles();
  std::vector<std::vector<uint32_t>> other(n), type(n);
  for (size_t k = 0; k < bonds_.size(); k += 3) {
    const uint32_t a = bonds_[k], b = bonds_[k + 1], t = bonds_[k + 2];
    other[a].push_back(b);
    type[a].push_back(t);
    other[b].push_back(a);
    type[b].push_back(t);
  }
  std::vector<uint32_t> start, flat_o, flat_t, s2;
  build_csr(n, other, 1, start, flat_o);
  build_csr(n, type, 1, s2, flat_t);
  const hipStream_t st = sample_->GetStream();
  upload_vec(bond_start_, start, st);
  upload_vec(bond_other_, flat_o, st);
  upload_vec(bond_type_, flat_t, st);
  if (slot_of_.size() != n) slot_of_.resize(n);
  slot_epoch_ = 0;
  sample_->InvalidateGraphs();
}

void MoleculeData::UploadAngles() {
  const unsigned n = sample_->GetNumberOfParticles();
  std::vector<std::vector<uint32_t>> per(n);
  for (size_t k = 0; k < angles_.size(); k += 4) {
    const uint32_t a = angles_[k], b = angles_[k + 1], c = angles_[k + 2];
    for (uint32_t o : {a, b, c}) per[o].insert(per[o].end(), angles_.begin() + k, angles_.begin() + k + 4);
  }
  std::vector<uint32_t> start, flat;
  build_csr(n, per, 4, start, flat);
  const hipStream_t st = sample_->GetStream();
  upload_vec(angle_start_, start, st);
  upload_vec(angle_atoms_, flat, st);
  if (slot_of_.size() != n) slot_of_.resize(n);
  slot_epoch_ = 0;
  sample_->InvalidateGraphs();
}

const uint32_t* MoleculeData::SlotIndex() {
  const unsigned n = sample_->GetNumberOfParticles();
  // allocated with the lists, never here: this runs inside captured graphs
  if (slot_of_.size() != n)
    throw RUMD_Error("MoleculeData", "SlotIndex", "the number of particles changed since the topology was set");
  // once per force call (the bonded potentials of a step share it; under a
  // graph capture this decides what is recorded)
  if (slot_epoch_ != sample_->GetSlotEpoch()) {
    rumdx_check(rumdx_slot_index(sample_->GetParticleData().Descriptor(), slot_of_.get(), sample_->GetStream()),
                "MoleculeData", "SlotIndex");
    slot_epoch_ = sample_->GetSlotEpoch();
  }
  return slot_of_.get();
}

// ------------------------------------------------------------------ BondHarmonic
// BondPotential::SetParams (src/BondPotential.cu:25-58): per bond type.
void BondHarmonic::SetParams(unsigned bond_type, float bond_length, float stiffness, bool exclude) {
  if (params_.size() < 2ull * (bond_type + 1)) {
    params_.resize(2ull * (bond_type + 1), 0.f);
    exclude_.resize(bond_type + 1, false);
  }
  params_[2 * bond_type] = bond_length;
  params_[2 * bond_type + 1] = stiffness;
  exclude_[bond_type] = exclude;
  d_params_.resize(params_.size());
  d_params_.upload(params_.data(), params_.size(), sample ? sample->GetStream() : nullptr);
  if (sample) sample->InvalidateGraphs();
}

std::vector<unsigned> BondHarmonic::GetExcludedTypes() const {
  std::vector<unsigned> t;
  for (unsigned k = 0; k < exclude_.size(); ++k)
    if (exclude_[k]) t.push_back(k);
  return t;
}

void BondHarmonic::SetExclusions(PairPotential* non_bond_pot) {
  if (!sample || sample->GetMoleculeData() == nullptr)
    throw RUMD_Error(GetID_String(), "SetExclusions", "no molecule data: read a topology first");
  const std::vector<uint32_t>& b = sample->GetMoleculeData()->GetBonds();
  std::vector<uint32_t> pairs;
  for (size_t k = 0; k < b.size(); k += 3)
    if (b[k + 2] < exclude_.size() && exclude_[b[k + 2]]) pairs.insert(pairs.end(), {b[k], b[k + 1]});
  if (!pairs.empty()) non_bond_pot->GetNeighborList().AddExclusions(pairs);
}

void BondHarmonic::Launch(const rumdx_particles* p, bool initialize, bool calc_stresses, const char* method) {
  if (!sample) throw RUMD_Error(GetID_String(), method, "no sample");
  MoleculeData* m = sample->GetMoleculeData();
  if (m == nullptr) throw RUMD_Error(GetID_String(), method, "no molecule data: read a topology first");
  if (params_.empty()) throw RUMD_Error(GetID_String(), method, "no bond parameters set");
  const uint32_t* slot_of = m->SlotIndex();
  const auto entry = fene_ ? rumdx_bond_fene : rumdx_bond_harmonic;
  rumdx_check(entry(p, slot_of, m->BondStart(), m->BondOther(), m->BondType(), d_params_.get(),
                    static_cast<uint32_t>(params_.size() / 2), sample->GetSimulationBox()->GetBox8(),
                    initialize ? 1 : 0, calc_stresses ? 1 : 0, sample->GetStream()),
              GetID_String().c_str(), method);
}

void BondHarmonic::CalcF(bool initialize, bool calc_stresses) {
  Launch(sample->GetParticleData().Descriptor(), initialize, calc_stresses, "CalcF");
}

double BondHarmonic::GetPotentialEnergy() {
  const unsigned nvp = sample->GetNumberOfVirtualParticles();
  if (pe_f_.size() != 4ull * nvp) {
    pe_f_.resize(4ull * nvp);
    pe_w_.resize(4ull * nvp);
  }
  rumdx_particles p = *sample->GetParticleData().Descriptor();
  p.f_dev = pe_f_.get();
  p.w_dev = pe_w_.get();
  Launch(&p, true, false, "GetPotentialEnergy");
  return sum_energy(sample, pe_f_.get(), "BondHarmonic");
}

// ------------------------------------------------------------------ AngleCosSq
// AnglePotential::SetParams (src/AnglePotential.cu:23-45): per angle type.
void AngleCosSq::SetParams(unsigned angle_type, float theta0, float ktheta) {
  if (params_.size() < 2ull * (angle_type + 1)) params_.resize(2ull * (angle_type + 1), 0.f);
  params_[2 * angle_type] = theta0;
  params_[2 * angle_type + 1] = ktheta;
  d_params_.resize(params_.size());
  d_params_.upload(params_.data(), params_.size(), sample ? sample->GetStream() : nullptr);
  if (sample) sample->InvalidateGraphs();
}

void AngleCosSq::SetExclusions(PairPotential* non_bond_pot) {
  if (!exclude_angle_) return;
  if (!sample || sample->GetMoleculeData() == nullptr)
    throw RUMD_Error(GetID_String(), "SetExclusions", "no molecule data: read a topology first");
  const std::vector<uint32_t>& a = sample->GetMoleculeData()->GetAngles();
  std::vector<uint32_t> pairs;
  for (size_t k = 0; k < a.size(); k += 4) pairs.insert(pairs.end(), {a[k], a[k + 2]});
  if (!pairs.empty()) non_bond_pot->GetNeighborList().AddExclusions(pairs);
}

void AngleCosSq::Launch(const rumdx_particles* p, bool initialize, bool calc_stresses, const char* method) {
  if (!sample) throw RUMD_Error(GetID_String(), method, "no sample");
  MoleculeData* m = sample->GetMoleculeData();
  if (m == nullptr) throw RUMD_Error(GetID_String(), method, "no molecule data: read a topology first");
  if (params_.empty()) throw RUMD_Error(GetID_String(), method, "no angle parameters set");
  if (m->GetNumberOfAngleTypes() > params_.size() / 2)
    throw RUMD_Error(GetID_String(), method, "an angle type of the topology has no parameters");
  const uint32_t* slot_of = m->SlotIndex();
  auto* entry = squared_ ? rumdx_angle_sq : rumdx_angle_cossq;
  rumdx_check(entry(p, slot_of, m->AngleStart(), m->AngleAtoms(), d_params_.get(),
                    static_cast<uint32_t>(params_.size() / 2), sample->GetSimulationBox()->GetBox8(),
                    initialize ? 1 : 0, calc_stresses ? 1 : 0, sample->GetStream()),
              GetID_String().c_str(), method);
}

void AngleCosSq::CalcF(bool initialize, bool calc_stresses) {
  Launch(sample->GetParticleData().Descriptor(), initialize, calc_stresses, "CalcF");
}

double AngleCosSq::GetPotentialEnergy() {
  const unsigned nvp = sample->GetNumberOfVirtualParticles();
  if (pe_f_.size() != 4ull * nvp) {
    pe_f_.resize(4ull * nvp);
    pe_w_.resize(4ull * nvp);
  }
  rumdx_particles p = *sample->GetParticleData().Descriptor();
  p.f_dev = pe_f_.get();
  p.w_dev = pe_w_.get();
  Launch(&p, true, false, "GetPotentialEnergy");
  return sum_energy(sample, pe_f_.get(), GetID_String().c_str());
}

// ------------------------------------------------------------------ DihedralRyckaert
// DihedralPotential::SetParams (src/DihedralPotential.cu:31-38): six
// coefficients per dihedral type.
void DihedralRyckaert::SetParams(unsigned dihedral_type, const std::vector<float>& coeffs) {
  if (coeffs.size() != num_params_)
    throw RUMD_Error(GetID_String(), "SetParams", "Wrong number of parameters in SetParams");
  if (params_.size() < 6ull * (dihedral_type + 1)) params_.resize(6ull * (dihedral_type + 1), 0.f);
  // six floats per type on the device, zero-padded (CopyParamsToGPU, src/DihedralPotential.cu:51-66)
  std::fill(params_.begin() + 6ull * dihedral_type, params_.begin() + 6ull * (dihedral_type + 1), 0.f);
  std::copy(coeffs.begin(), coeffs.end(), params_.begin() + 6ull * dihedral_type);
  d_params_.resize(params_.size());
  d_params_.upload(params_.data(), params_.size(), sample ? sample->GetStream() : nullptr);
  if (sample) sample->InvalidateGraphs();
}

void DihedralRyckaert::SetExclusions(PairPotential* non_bond_pot) {
  if (!exclude_dihedral_) return;
  if (!sample || sample->GetMoleculeData() == nullptr)
    throw RUMD_Error(GetID_String(), "SetExclusions", "no molecule data: read a topology first");
  const std::vector<uint32_t>& d = sample->GetMoleculeData()->GetDihedrals();
  std::vector<uint32_t> pairs;
  for (size_t k = 0; k < d.size(); k += 5) pairs.insert(pairs.end(), {d[k], d[k + 3]});
  if (!pairs.empty()) non_bond_pot->GetNeighborList().AddExclusions(pairs);
}

void DihedralRyckaert::Launch(const rumdx_particles* p, bool initialize, bool calc_stresses, const char* method) {
  if (!sample) throw RUMD_Error(GetID_String(), method, "no sample");
  MoleculeData* m = sample->GetMoleculeData();
  if (m == nullptr) throw RUMD_Error(GetID_String(), method, "no molecule data: read a topology first");
  if (params_.empty()) throw RUMD_Error(GetID_String(), method, "no dihedral parameters set");
  if (m->GetNumberOfDihedralTypes() > params_.size() / 6)
    throw RUMD_Error(GetID_String(), method, "a dihedral type of the topology has no parameters");
  const uint32_t* slot_of = m->SlotIndex();
  auto* entry = periodic_ ? rumdx_dihedral_periodic : rumdx_dihedral_ryckaert;
  rumdx_check(entry(p, slot_of, m->DihedralStart(), m->DihedralAtoms(), d_params_.get(),
                    static_cast<uint32_t>(params_.size() / 6), sample->GetSimulationBox()->GetBox8(),
                    initialize ? 1 : 0, calc_stresses ? 1 : 0, sample->GetStream()),
              GetID_String().c_str(), method);
}

void DihedralRyckaert::CalcF(bool initialize, bool calc_stresses) {
  Launch(sample->GetParticleData().Descriptor(), initialize, calc_stresses, "CalcF");
}

double DihedralRyckaert::GetPotentialEnergy() {
  const unsigned nvp = sample->GetNumberOfVirtualParticles();
  if (pe_f_.size() != 4ull * nvp) {
    pe_f_.resize(4ull * nvp);
    pe_w_.resize(4ull * nvp);
  }
  rumdx_particles p = *sample->GetParticleData().Descriptor();
  p.f_dev = pe_f_.get();
  p.w_dev = pe_w_.get();
  Launch(&p, true, false, "GetPotentialEnergy");
  return sum_energy(sample, pe_f_.get(), GetID_String().c_str());
}
