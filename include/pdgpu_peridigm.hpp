// Host-side C++ mirror of Peridigm's material / damage-model interface over
// the pdgpu C ABI (header only).
//
// pdgpu::ElasticMaterial and pdgpu::CriticalStretchDamageModel keep the
// calling convention of PeridigmNS::Material::initialize / computeForce
// (src/materials/Peridigm_Material.hpp:120-141) and
// PeridigmNS::DamageModel::computeDamage (src/damage/Peridigm_DamageModel.hpp:106-111):
//   (dt, numOwnedPoints, ownedIDs, neighborhoodList, DataManager&)
// and read/write the same fields (Model_Coordinates, Coordinates, Volume,
// Weighted_Volume, Dilatation, Force_Density, Bond_Damage, Damage) with the
// same N / NP1 / NONE states. pdgpu::DataManager is a minimal stand-in for
// PeridigmNS::DataManager (src/core/Peridigm_DataManager.hpp:66-256): flat
// FP64 arrays by field name and step, xyz-interleaved for vectors, bond
// fields in neighbourhood-list order, with the N <-> NP1 swap of updateState.
// Errors are thrown as std::runtime_error, as the reference throws through
// TEUCHOS_TEST_FOR_EXCEPT_MSG.
#ifndef PDGPU_PERIDIGM_HPP
#define PDGPU_PERIDIGM_HPP

#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "pdgpu.h"

namespace pdgpu {

enum Step { STEP_NONE = 0, STEP_N = 1, STEP_NP1 = 2 };

class DataManager {
 public:
  DataManager(int numPoints, long long numBonds) : n_(numPoints), b_(numBonds) {}
  int numPoints() const { return n_; }
  long long numBonds() const { return b_; }

  // Allocates (zeroed) on first use: node vectors 3N, scalars N, bond fields B.
  double* getData(const std::string& field, Step step) {
    auto key = std::make_pair(field, step);
    auto it = data_.find(key);
    if (it == data_.end()) it = data_.emplace(key, std::vector<double>(size_of(field), 0.0)).first;
    return it->second.data();
  }
  // Swap N and NP1 of every two-step field (DataManager::updateState).
  void updateState() {
    for (auto& kv : data_) {
      if (kv.first.second != STEP_N) continue;
      auto np1 = data_.find(std::make_pair(kv.first.first, STEP_NP1));
      if (np1 != data_.end()) kv.second.swap(np1->second);
    }
  }

 private:
  size_t size_of(const std::string& f) const {
    if (f == "Model_Coordinates" || f == "Coordinates" || f == "Force_Density" ||
        f == "Displacement" || f == "Velocity" || f == "Acceleration")
      return 3 * (size_t)n_;
    if (f == "Bond_Damage") return (size_t)b_;
    return (size_t)n_;
  }
  int n_;
  long long b_;
  std::map<std::pair<std::string, Step>, std::vector<double>> data_;
};

namespace detail {

inline void check(pdg_ctx* c, int rc, const char* what) {
  if (rc != PDG_OK)
    throw std::runtime_error(std::string("**** Error in ") + what + ": " +
                             (c ? pdg_last_error(c) : "no context"));
}

inline long long list_size(int numOwned, const int* nlist) {
  long long li = 0;
  for (int i = 0; i < numOwned; i++) li += nlist[li] + 1;
  return li;
}

// One device context per neighbourhood list, shared by the material and the
// damage model of a block (one CSR build and one set of device arrays). The
// material sets the elastic constants and influence function, the damage
// model the critical stretch; the context is rebuilt only when the list,
// the coordinates or those constants change. Bindings are recognised by the
// arrays' addresses and sizes, as Peridigm keeps them for a block's life;
// after editing a bound array in place, call invalidate().
class Device {
 public:
  Device() {
    int rc = pdg_ctx_create(0, &ctx_);
    if (rc != PDG_OK) throw std::runtime_error("pdgpu: no HIP device");
  }
  ~Device() { pdg_ctx_destroy(ctx_); }
  Device(const Device&) = delete;
  Device& operator=(const Device&) = delete;
  pdg_ctx* get() const { return ctx_; }

  void bind_material(int numOwned, const int* ownedIDs, const int* nlist, const double* x,
                     const double* vol, double K, double mu, double horizon, int influence) {
    Consts c = c_;
    c.K = K; c.mu = mu; c.horizon = horizon; c.influence = influence; c.material = true;
    bind(numOwned, ownedIDs, nlist, x, vol, c);
  }
  void bind_damage(int numOwned, const int* ownedIDs, const int* nlist, const double* x,
                   const double* vol, double horizon, double sc) {
    Consts c = c_;
    if (!c.material) c.horizon = horizon;
    c.sc = sc;
    bind(numOwned, ownedIDs, nlist, x, vol, c);
  }
  void invalidate() { n_ = -1; }

 private:
  struct Consts {
    double K = 1.0, mu = 1.0, horizon = 0.0, sc = 0.0;
    int influence = PDG_INFLUENCE_ONE;
    bool material = false;
    bool operator==(const Consts& o) const {
      return K == o.K && mu == o.mu && horizon == o.horizon && sc == o.sc &&
             influence == o.influence && material == o.material;
    }
  };
  void bind(int numOwned, const int* ownedIDs, const int* nlist, const double* x,
            const double* vol, const Consts& c) {
    const bool same_list = nlist == nlist_ && numOwned == n_ && x == x_ && vol == vol_;
    if (same_list && c == c_) return;
    if (!same_list) {
      size_ = list_size(numOwned, nlist);
      check(ctx_, pdg_set_topology(ctx_, numOwned, ownedIDs, nlist, size_), "pdg_set_topology");
    }
    check(ctx_, pdg_set_constants(ctx_, x, vol, c.K, c.mu, c.horizon, c.influence, c.sc),
          "pdg_set_constants");
    nlist_ = nlist; n_ = numOwned; x_ = x; vol_ = vol; c_ = c;
  }
  pdg_ctx* ctx_ = nullptr;
  const int* nlist_ = nullptr;
  const double* x_ = nullptr;
  const double* vol_ = nullptr;
  long long size_ = 0;
  int n_ = -1;
  Consts c_;
};

// The context shared by every model bound to the same neighbourhood list
// (held while any of them lives).
inline std::shared_ptr<Device> shared_device(const int* nlist) {
  static std::map<const int*, std::weak_ptr<Device>> reg;
  auto it = reg.find(nlist);
  if (it != reg.end())
    if (auto d = it->second.lock()) return d;
  auto d = std::make_shared<Device>();
  reg[nlist] = d;
  return d;
}

}  // namespace detail

struct ElasticParams {
  double density = 0.0, bulkModulus = 0.0, shearModulus = 0.0, horizon = 0.0;
  int influence = PDG_INFLUENCE_ONE;
};

// Linear peridynamic solid ("Elastic"), src/materials/Peridigm_ElasticMaterial.cpp:59-172.
class ElasticMaterial {
 public:
  explicit ElasticMaterial(const ElasticParams& p) : p_(p) {}
  double Density() const { return p_.density; }
  double BulkModulus() const { return p_.bulkModulus; }
  double ShearModulus() const { return p_.shearModulus; }

  // Weighted_Volume (STEP_NONE) from Model_Coordinates and Volume.
  void initialize(double /*dt*/, int numOwnedPoints, const int* ownedIDs,
                  const int* neighborhoodList, DataManager& dm) const {
    bind(numOwnedPoints, ownedIDs, neighborhoodList, dm);
    detail::check(dev_->get(), pdg_initialize(dev_->get(), dm.getData("Weighted_Volume", STEP_NONE)),
                  "pdg_initialize");
  }
  // After editing a bound neighbourhood list or coordinates in place.
  void invalidate() const { if (dev_) dev_->invalidate(); }
  // The device context in use (shared with a damage model on the same list).
  const pdg_ctx* context() const { return dev_ ? dev_->get() : nullptr; }
  // Force_Density and Dilatation (STEP_NP1) from Coordinates NP1 and
  // Bond_Damage NP1; Force_Density is overwritten (zeroed first in the reference).
  void computeForce(double dt, int numOwnedPoints, const int* ownedIDs,
                    const int* neighborhoodList, DataManager& dm) const {
    bind(numOwnedPoints, ownedIDs, neighborhoodList, dm);
    detail::check(dev_->get(),
                  pdg_compute_force(dev_->get(), dt, dm.getData("Coordinates", STEP_NP1),
                                    dm.getData("Bond_Damage", STEP_NP1),
                                    dm.getData("Dilatation", STEP_NP1),
                                    dm.getData("Force_Density", STEP_NP1)),
                  "pdg_compute_force");
  }

 private:
  void bind(int n, const int* ids, const int* nl, DataManager& dm) const {
    if (!dev_ || nl != list_) { dev_ = detail::shared_device(nl); list_ = nl; }
    dev_->bind_material(n, ids, nl, dm.getData("Model_Coordinates", STEP_NONE),
                        dm.getData("Volume", STEP_NONE), p_.bulkModulus, p_.shearModulus,
                        p_.horizon, p_.influence);
  }
  ElasticParams p_;
  mutable std::shared_ptr<detail::Device> dev_;
  mutable const int* list_ = nullptr;
};

// Critical stretch, src/damage/Peridigm_CriticalStretchDamageModel.cpp:55-184.
class CriticalStretchDamageModel {
 public:
  CriticalStretchDamageModel(double criticalStretch, double horizon)
      : sc_(criticalStretch), horizon_(horizon) {}
  // Damage and Bond_Damage (NP1) set to zero.
  void initialize(double /*dt*/, int numOwnedPoints, const int* /*ownedIDs*/,
                  const int* /*neighborhoodList*/, DataManager& dm) const {
    double* d = dm.getData("Damage", STEP_NP1);
    for (int i = 0; i < numOwnedPoints; i++) d[i] = 0.0;
    double* bd = dm.getData("Bond_Damage", STEP_NP1);
    for (long long b = 0; b < dm.numBonds(); b++) bd[b] = 0.0;
  }
  // Bond_Damage NP1 from Bond_Damage N and Coordinates NP1; Damage NP1.
  void computeDamage(double /*dt*/, int numOwnedPoints, const int* ownedIDs,
                     const int* neighborhoodList, DataManager& dm) const {
    if (!dev_ || neighborhoodList != list_) {
      dev_ = detail::shared_device(neighborhoodList);
      list_ = neighborhoodList;
    }
    dev_->bind_damage(numOwnedPoints, ownedIDs, neighborhoodList,
                      dm.getData("Model_Coordinates", STEP_NONE), dm.getData("Volume", STEP_NONE),
                      horizon_, sc_);
    long long near = 0;
    detail::check(dev_->get(),
                  pdg_compute_damage(dev_->get(), dm.getData("Coordinates", STEP_NP1),
                                     dm.getData("Bond_Damage", STEP_N),
                                     dm.getData("Bond_Damage", STEP_NP1),
                                     dm.getData("Damage", STEP_NP1), &near),
                  "pdg_compute_damage");
    nearCritical_ = near;
  }
  long long nearCriticalBonds() const { return nearCritical_; }
  const pdg_ctx* context() const { return dev_ ? dev_->get() : nullptr; }
  void invalidate() const { if (dev_) dev_->invalidate(); }

 private:
  double sc_, horizon_;
  mutable std::shared_ptr<detail::Device> dev_;
  mutable const int* list_ = nullptr;
  mutable long long nearCritical_ = 0;
};

}  // namespace pdgpu

#endif  // PDGPU_PERIDIGM_HPP
