// Field storage and layout rules of the host mirror.
#include "core.h"

#include <algorithm>
#include <numeric>

namespace Kripke {
namespace Core {

int default_order_position(int layout, Idx t) {
  // DefaultOrder<Layout>::type, src/Kripke/VarTypes.h:68-101 (entry 0 is `long`).
  using I = Idx;
  static const std::vector<I> lists[6] = {
      {I::Dimension, I::Material, I::Direction, I::Legendre, I::Moment, I::GlobalGroup, I::Group, I::Zone, I::ZoneK,
       I::ZoneJ, I::ZoneI, I::MixElem},  // DGZ
      {I::Dimension, I::Material, I::Direction, I::Legendre, I::Moment, I::Zone, I::ZoneK, I::ZoneJ, I::ZoneI,
       I::GlobalGroup, I::Group, I::MixElem},  // DZG
      {I::Dimension, I::Material, I::GlobalGroup, I::Group, I::Direction, I::Legendre, I::Moment, I::Zone, I::ZoneK,
       I::ZoneJ, I::ZoneI, I::MixElem},  // GDZ
      {I::Dimension, I::Material, I::GlobalGroup, I::Group, I::Zone, I::ZoneK, I::ZoneJ, I::ZoneI, I::MixElem,
       I::Direction, I::Legendre, I::Moment},  // GZD
      {I::Dimension, I::Material, I::Zone, I::ZoneK, I::ZoneJ, I::ZoneI, I::MixElem, I::Direction, I::Legendre,
       I::Moment, I::GlobalGroup, I::Group},  // ZDG
      {I::Dimension, I::Material, I::Zone, I::ZoneK, I::ZoneJ, I::ZoneI, I::MixElem, I::GlobalGroup, I::Group,
       I::Direction, I::Legendre, I::Moment},  // ZGD
  };
  const auto &l = lists[layout];
  for (size_t k = 0; k < l.size(); ++k)
    if (l[k] == t) return (int)k + 1;
  return -1;
}

FieldBase::FieldBase(kr_context *ctx, const Set &set, std::vector<Idx> idx, int layout, size_t elem, bool host,
                     bool device)
    : m_ctx(ctx), m_set(&set), m_elem(elem) {
  if (idx.size() != set.getNumDimensions())
    throw std::domain_error("Number of dimensions must match between Field and Set");  // Field.h:190
  initChunks(set);
  const size_t nd = idx.size();
  // Permutation: argument positions sorted by DefaultOrder position (stable
  // for repeated index types), VarLayout.h:87-135.
  std::vector<size_t> perm(nd);
  std::iota(perm.begin(), perm.end(), 0);
  std::stable_sort(perm.begin(), perm.end(), [&](size_t a, size_t b) {
    return default_order_position(layout, idx[a]) < default_order_position(layout, idx[b]);
  });
  m_chunk_size.resize(numChunks());
  m_chunk_off.resize(numChunks());
  m_strides.resize(numChunks());
  size_t off = 0;
  for (size_t c = 0; c < numChunks(); ++c) {
    SdomId s = chunkSdom(c);
    std::vector<int64_t> ext(nd), st(nd);
    for (size_t d = 0; d < nd; ++d) ext[d] = (int64_t)set.dimSize(s, d);
    int64_t acc = 1;
    for (size_t k = nd; k-- > 0;) {
      st[perm[k]] = acc;
      acc *= ext[perm[k]];
    }
    m_strides[c] = st;
    m_chunk_size[c] = (size_t)acc;
    m_chunk_off[c] = off;
    // keep chunks 256-byte aligned
    off += ((size_t)acc * elem + 255) / 256 * 256 / elem;
  }
  m_total = off;
  if (host) m_host.assign(m_total * m_elem, 0);
  if (device && ctx) {
    if (kr_malloc(ctx, std::max<size_t>(m_total * m_elem, 8), &m_dev) != KR_OK)
      throw std::runtime_error(std::string("field allocation failed: ") + kr_last_error());
    // chunks are padded to 256 bytes, so the slab is a whole number of doubles
    if (kr_kconst(ctx, (double *)m_dev, (int64_t)(m_total * m_elem / sizeof(double)), 0.0) != KR_OK)
      throw std::runtime_error(std::string("field zero-fill failed: ") + kr_last_error());
  }
}

FieldBase::~FieldBase() {
  if (m_dev) kr_free(m_ctx, m_dev);
}

int FieldBase::upload() {
  if (!m_dev || m_host.empty()) return KR_OK;
  return kr_memcpy_htod(m_ctx, m_dev, m_host.data(), m_total * m_elem);
}

}  // namespace Core
}  // namespace Kripke
