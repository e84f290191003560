// Host mirror of Kripke's core data model, written from scratch:
// DataStore (src/Kripke/Core/DataStore.h:23-88), PartitionSpace
// (src/Kripke/Core/PartitionSpace.h:35-97), the Set family
// (src/Kripke/Core/Set.h:23-202) and Field (src/Kripke/Core/Field.h:29-290).
// Field chunks live in HBM (one allocation per field, one chunk per
// subdomain of the field's space); small setup fields also keep a host copy.
#pragma once
#include <stdint.h>

#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../../include/kripke_amd.h"

namespace Kripke {

using SdomId = int32_t;

namespace Core {

class BaseVar {
 public:
  virtual ~BaseVar() = default;
};

// Named variable store; the store owns what is added to it.
class DataStore {
 public:
  DataStore() = default;
  DataStore(const DataStore &) = delete;
  DataStore &operator=(const DataStore &) = delete;
  ~DataStore() {
    // delete in reverse order of insertion: later variables may refer to earlier ones
    for (auto it = m_order.rbegin(); it != m_order.rend(); ++it) {
      auto f = m_vars.find(*it);
      if (f != m_vars.end()) delete f->second;
    }
  }
  void addVariable(const std::string &name, BaseVar *var) {
    if (m_vars.count(name)) throw std::domain_error("Variable '" + name + "' already exists");
    m_vars[name] = var;
    m_order.push_back(name);
  }
  template <class T, class... Args>
  T &newVariable(const std::string &name, Args &&...args) {
    T *v = new T(std::forward<Args>(args)...);
    addVariable(name, v);
    return *v;
  }
  template <class T>
  T &getVariable(const std::string &name) const {
    auto it = m_vars.find(name);
    if (it == m_vars.end()) throw std::domain_error("Cannot find '" + name + "' in DataStore");
    T *v = dynamic_cast<T *>(it->second);
    if (!v) throw std::domain_error("Variable '" + name + "' has a different type");
    return *v;
  }
  bool isVariableDefined(const std::string &name) const { return m_vars.count(name) != 0; }
  template <class T>
  bool isVariableType(const std::string &name) const {
    auto it = m_vars.find(name);
    return it != m_vars.end() && dynamic_cast<T *>(it->second) != nullptr;
  }
  void deleteVariable(const std::string &name) {
    auto it = m_vars.find(name);
    if (it == m_vars.end()) throw std::domain_error("Cannot find '" + name + "' in DataStore");
    delete it->second;
    m_vars.erase(it);
  }
  std::vector<std::string> getVariableList() const {
    std::vector<std::string> v;
    for (auto &kv : m_vars) v.push_back(kv.first);
    return v;
  }

 private:
  std::map<std::string, BaseVar *> m_vars;
  std::vector<std::string> m_order;
};

enum SPACE { SPACE_P = 0, SPACE_Q, SPACE_RX, SPACE_RY, SPACE_RZ, SPACE_R, SPACE_PR, SPACE_PQR, SPACE_NULL, NUM_SPACES };

// On-rank phase-space decomposition: subdomains laid out row-major over
// (P, Q, Rx, Ry, Rz) (PartitionSpace.cpp:75-123); a space's chunk id is the
// row-major index over the dimensions that space keeps.
class PartitionSpace : public BaseVar {
 public:
  PartitionSpace(int SP, int SQ, int Sx, int Sy, int Sz) : m_n{SP, SQ, Sx, Sy, Sz} {}
  int dim(int d) const { return m_n[d]; }
  size_t getNumSubdomains(SPACE space = SPACE_PQR) const {
    size_t n = 1;
    for (int d = 0; d < 5; ++d)
      if (keeps(space, d)) n *= (size_t)m_n[d];
    return n;
  }
  void sdomIdToCoord(SdomId s, int c[5]) const {
    int rem = s;
    for (int d = 4; d >= 0; --d) {
      c[d] = rem % m_n[d];
      rem /= m_n[d];
    }
  }
  SdomId coordToSdomId(const int c[5]) const {
    int id = 0;
    for (int d = 0; d < 5; ++d) id = id * m_n[d] + c[d];
    return id;
  }
  size_t subdomainToSpace(SPACE space, SdomId s) const {
    int c[5];
    sdomIdToCoord(s, c);
    size_t id = 0;
    for (int d = 0; d < 5; ++d)
      if (keeps(space, d)) id = id * (size_t)m_n[d] + (size_t)c[d];
    return id;
  }
  SdomId spaceToSubdomain(SPACE space, size_t chunk) const {
    int c[5] = {0, 0, 0, 0, 0};
    size_t rem = chunk;
    for (int d = 4; d >= 0; --d)
      if (keeps(space, d)) {
        c[d] = (int)(rem % (size_t)m_n[d]);
        rem /= (size_t)m_n[d];
      }
    return coordToSdomId(c);
  }
  static bool keeps(SPACE space, int d) {
    switch (space) {
      case SPACE_P: return d == 0;
      case SPACE_Q: return d == 1;
      case SPACE_RX: return d == 2;
      case SPACE_RY: return d == 3;
      case SPACE_RZ: return d == 4;
      case SPACE_R: return d >= 2;
      case SPACE_PR: return d != 1;
      case SPACE_PQR: return true;
      default: return false;
    }
  }

 private:
  int m_n[5];
};

// Chunked decomposition shared by Sets and Fields (DomainVar.cpp:15-47).
class DomainVar : public BaseVar {
 public:
  size_t getNumSubdomains() const { return m_sdom_to_chunk.size(); }
  const std::vector<SdomId> &getWorkList() const { return m_work_list; }
  size_t chunkOf(SdomId s) const { return m_sdom_to_chunk.at((size_t)s); }
  size_t numChunks() const { return m_chunk_to_sdom.size(); }
  SdomId chunkSdom(size_t c) const { return m_chunk_to_sdom.at(c); }

 protected:
  void initChunks(const PartitionSpace &ps, SPACE space) {
    size_t ns = ps.getNumSubdomains(SPACE_PQR), nc = ps.getNumSubdomains(space);
    m_sdom_to_chunk.resize(ns);
    for (size_t s = 0; s < ns; ++s) m_sdom_to_chunk[s] = ps.subdomainToSpace(space, (SdomId)s);
    m_chunk_to_sdom.resize(nc);
    m_work_list.resize(nc);
    for (size_t c = 0; c < nc; ++c) m_chunk_to_sdom[c] = m_work_list[c] = ps.spaceToSubdomain(space, c);
  }
  void initChunks(const DomainVar &o) {
    m_sdom_to_chunk = o.m_sdom_to_chunk;
    m_chunk_to_sdom = o.m_chunk_to_sdom;
    m_work_list = o.m_work_list;
  }
  std::vector<size_t> m_sdom_to_chunk;
  std::vector<SdomId> m_chunk_to_sdom;
  std::vector<SdomId> m_work_list;
};

class Set : public DomainVar {
 public:
  size_t size(SdomId s) const { return m_chunk_size.at(chunkOf(s)); }
  size_t lower(SdomId s) const { return m_chunk_lower.at(chunkOf(s)); }
  size_t globalSize() const { return m_global_size; }
  virtual size_t getNumDimensions() const = 0;
  virtual size_t dimSize(SdomId s, size_t) const { return size(s); }

 protected:
  std::vector<size_t> m_chunk_size, m_chunk_lower;
  size_t m_global_size = 0;
};

// Set.cpp:34-80
class RangeSet : public Set {
 public:
  RangeSet(const PartitionSpace &ps, SPACE space, const std::vector<size_t> &local) {
    initChunks(ps, space);
    if (local.size() != numChunks()) throw std::domain_error("RangeSet: wrong number of local sizes");
    m_chunk_size = local;
    m_chunk_lower.resize(local.size());
    size_t acc = 0;
    for (size_t c = 0; c < local.size(); ++c) {
      m_chunk_lower[c] = acc;
      acc += local[c];
    }
    m_global_size = acc;
  }
  size_t getNumDimensions() const override { return 1; }
};

// Set.cpp:114-147
class GlobalRangeSet : public Set {
 public:
  GlobalRangeSet(const PartitionSpace &ps, size_t global_size) {
    initChunks(ps, SPACE_NULL);
    m_chunk_size = {global_size};
    m_chunk_lower = {0};
    m_global_size = global_size;
  }
  size_t getNumDimensions() const override { return 1; }
};

// ProductSet<N> (Set.h:127-202): chunk sizes are products of the spanned sets.
class ProductSet : public Set {
 public:
  ProductSet(const PartitionSpace &ps, SPACE space, std::vector<const Set *> spanned) : m_spanned(spanned) {
    initChunks(ps, space);
    m_chunk_size.assign(numChunks(), 1);
    m_chunk_lower.assign(numChunks(), 0);
    for (size_t c = 0; c < numChunks(); ++c)
      for (auto *s : m_spanned) m_chunk_size[c] *= s->size(chunkSdom(c));
    m_global_size = 1;
    for (auto *s : m_spanned) m_global_size *= s->globalSize();
  }
  size_t getNumDimensions() const override { return m_spanned.size(); }
  size_t dimSize(SdomId s, size_t d) const override { return m_spanned.at(d)->size(s); }

 private:
  std::vector<const Set *> m_spanned;
};

// Index types of Field<T, IDX...> (src/Kripke/VarTypes.h:18-30).
enum class Idx { Dimension, Material, Direction, Legendre, Moment, GlobalGroup, Group, Zone, ZoneK, ZoneJ, ZoneI, MixElem };

// Position of each index type in DefaultOrder<layout> (VarTypes.h:68-101);
// lower = longer stride.
int default_order_position(int layout, Idx t);

// Field storage: one HBM allocation, one chunk per subdomain of the set's
// space; strides from the DefaultOrder permutation (VarLayout.h:87-135).
class FieldBase : public DomainVar {
 public:
  virtual ~FieldBase();
  size_t size(SdomId s) const { return m_chunk_size.at(chunkOf(s)); }
  const Set &getSet() const { return *m_set; }
  size_t elemBytes() const { return m_elem; }
  bool onDevice() const { return m_dev != nullptr; }
  bool onHost() const { return !m_host.empty(); }
  void *deviceChunk(SdomId s) const { return m_dev ? (char *)m_dev + m_chunk_off.at(chunkOf(s)) * m_elem : nullptr; }
  void *hostChunk(SdomId s) { return m_host.empty() ? nullptr : m_host.data() + m_chunk_off.at(chunkOf(s)) * m_elem; }
  const std::vector<int64_t> &strides(SdomId s) const { return m_strides.at(chunkOf(s)); }
  size_t totalElements() const { return m_total; }
  // copy the whole host image to HBM
  int upload();
  void *deviceBase() const { return m_dev; }
  kr_context *context() const { return m_ctx; }

 protected:
  FieldBase(kr_context *ctx, const Set &set, std::vector<Idx> idx, int layout, size_t elem, bool host, bool device);
  kr_context *m_ctx;
  const Set *m_set;
  size_t m_elem;
  std::vector<size_t> m_chunk_size, m_chunk_off;
  std::vector<std::vector<int64_t>> m_strides;
  size_t m_total = 0;
  std::vector<char> m_host;
  void *m_dev = nullptr;
};

template <class T>
class Field : public FieldBase {
 public:
  using ElementType = T;
  Field(kr_context *ctx, const Set &set, std::vector<Idx> idx, int layout, bool host, bool device)
      : FieldBase(ctx, set, idx, layout, sizeof(T), host, device) {}
  T *host(SdomId s) { return (T *)hostChunk(s); }
  T *device(SdomId s) const { return (T *)deviceChunk(s); }
  // host element by indices in the field's own argument order
  T &at(SdomId s, std::initializer_list<int64_t> ix) {
    const auto &st = strides(s);
    int64_t off = 0, k = 0;
    for (int64_t v : ix) off += v * st[k++];
    return host(s)[off];
  }
};

}  // namespace Core
}  // namespace Kripke
