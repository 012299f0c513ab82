// tree_host.cpp -- host-side TreeMesh adaptation and connectivity, native.
//
// The tree is kept as its leaves in element (depth-first) order: level,
// integer coordinates at that level, centre, and the origin of each leaf in the
// current adaptation (kind 0 = old leaf src, 1 + q = child q of old leaf src,
// -1 = parent of old leaves src .. src + 2^D - 1). Every cell of the tree
// (leaves and their ancestors) is found by key (level, coordinates) in an
// open-addressing hash table; a leaf maps to its element index, an inner cell
// to -1.
//
//   refine_unbalanced!            abstract_tree.jl:317-353 (children in place of
//                                 the parent, child bit d = upper half in d)
//   refine! + rebalance!          abstract_tree.jl:367-393, 453-490 (the coarse
//                                 face neighbour of a refined cell without a
//                                 same-level neighbour is refined, until none is)
//   coarsen! (AMR selection)      amr.jl:304-370, abstract_tree.jl:516-600
//                                 (parents whose 2^D children all ask for it,
//                                 unless a child has a non-sibling face
//                                 neighbour with children; highest level first)
//   init_interfaces!/init_mortars!/init_boundaries!
//                                 containers_2d.jl:228-290, 475-543, 707-800;
//                                 containers_3d.jl:229-279, 391-480, 652-788
#include <algorithm>
#include <cstdint>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/trixi_hip.h"

struct trixi_hip_tree {
  int D = 0;
  int per[3] = {0, 0, 0};
  double length0 = 0;
  // leaves
  std::vector<int32_t> lv;
  std::vector<int64_t> ic;  // [n][D]
  std::vector<double> ctr;  // [n][D]
  std::vector<int32_t> kind, src;
  // connectivity of the current leaves (valid when connected)
  bool connected = false;
  std::vector<int32_t> nbr;            // [n][2D]
  std::vector<int32_t> if_ids, if_or;  // [I][2], [I]
  std::vector<int32_t> m_ids, m_side, m_or;  // [M][S + 1], [M], [M]
  std::vector<int32_t> b_ids, b_or, b_side;  // [B]
  std::string err;
  size_t n() const { return lv.size(); }
};

namespace {

constexpr int kShift[3] = {6, 25, 44};  // level in bits 0-5, coordinate d in 19 bits from kShift[d]
constexpr int kMaxLevel = 19;

inline uint64_t cell_key(int level, const int64_t *c, int D) {
  uint64_t k = (uint64_t)level;
  for (int d = 0; d < D; ++d) k |= (uint64_t)c[d] << kShift[d];
  return k;
}

// Cells of the tree by key: leaf -> element index, inner cell -> -1.
class CellMap {
 public:
  explicit CellMap(const trixi_hip_tree &t) {
    const int D = t.D;
    size_t cap = 64;
    while (cap < 4 * t.n() + 64) cap <<= 1;
    keys_.assign(cap, kEmpty);
    vals_.assign(cap, 0);
    mask_ = cap - 1;
    for (size_t i = 0; i < t.n(); ++i) put(cell_key(t.lv[i], &t.ic[i * D], D), (int32_t)i);
    int64_t c[3];
    for (size_t i = 0; i < t.n(); ++i) {
      for (int l = t.lv[i] - 1; l >= 0; --l) {  // ancestors up to the first one already present
        for (int d = 0; d < D; ++d) c[d] = t.ic[i * D + d] >> (t.lv[i] - l);
        if (!put(cell_key(l, c, D), -1)) break;
      }
    }
  }
  // (found, value)
  bool find(uint64_t k, int32_t *v) const {
    for (uint64_t s = hash(k) & mask_;; s = (s + 1) & mask_) {
      if (keys_[s] == k) {
        *v = vals_[s];
        return true;
      }
      if (keys_[s] == kEmpty) return false;
    }
  }

 private:
  static constexpr uint64_t kEmpty = ~(uint64_t)0;
  static uint64_t hash(uint64_t k) {  // splitmix64 finaliser
    k ^= k >> 30;
    k *= 0xbf58476d1ce4e5b9ULL;
    k ^= k >> 27;
    k *= 0x94d049bb133111ebULL;
    k ^= k >> 31;
    return k;
  }
  // inserts k -> v unless k is present; true if inserted
  bool put(uint64_t k, int32_t v) {
    for (uint64_t s = hash(k) & mask_;; s = (s + 1) & mask_) {
      if (keys_[s] == k) return false;
      if (keys_[s] == kEmpty) {
        keys_[s] = k;
        vals_[s] = v;
        return true;
      }
    }
  }
  std::vector<uint64_t> keys_;
  std::vector<int32_t> vals_;
  uint64_t mask_ = 0;
};

// Same-level cell across face dirn (2o = -o, 2o + 1 = +o), periodic dimensions
// wrapped; false outside a non-periodic domain.
inline bool shift(const trixi_hip_tree &t, int level, const int64_t *c, int dirn, int64_t *nb) {
  const int o = dirn / 2;
  const int64_t nside = (int64_t)1 << level;
  const int64_t k = c[o] + (dirn % 2 ? 1 : -1);
  for (int d = 0; d < t.D; ++d) nb[d] = c[d];
  nb[o] = k & (nside - 1);  // k in [-1, nside]: periodic wrap
  return t.per[o] || (k >= 0 && k < nside);
}

// Each marked leaf replaced in place by its 2^D children.
void refine_unbalanced(trixi_hip_tree &t, const std::vector<uint8_t> &marked) {
  const int D = t.D, nch = 1 << D;
  size_t cnt = 0;
  for (size_t i = 0; i < t.n(); ++i) cnt += marked[i] ? nch : 1;
  std::vector<int32_t> lv, kind, src;
  std::vector<int64_t> ic;
  std::vector<double> ctr;
  lv.reserve(cnt);
  kind.reserve(cnt);
  src.reserve(cnt);
  ic.reserve(cnt * D);
  ctr.reserve(cnt * D);
  for (size_t i = 0; i < t.n(); ++i) {
    if (!marked[i]) {
      lv.push_back(t.lv[i]);
      kind.push_back(t.kind[i]);
      src.push_back(t.src[i]);
      for (int d = 0; d < D; ++d) {
        ic.push_back(t.ic[i * D + d]);
        ctr.push_back(t.ctr[i * D + d]);
      }
      continue;
    }
    const double child_length = t.length0 / (double)((int64_t)1 << t.lv[i]) / 2;
    for (int q = 0; q < nch; ++q) {
      lv.push_back(t.lv[i] + 1);
      kind.push_back(1 + q);
      src.push_back(t.src[i]);
      for (int d = 0; d < D; ++d) {
        const int bit = (q >> d) & 1;
        ic.push_back(2 * t.ic[i * D + d] + bit);
        ctr.push_back(t.ctr[i * D + d] + (bit ? 1.0 : -1.0) * (child_length / 2));
      }
    }
  }
  t.lv.swap(lv);
  t.kind.swap(kind);
  t.src.swap(src);
  t.ic.swap(ic);
  t.ctr.swap(ctr);
  t.connected = false;
}

int64_t refine_balanced(trixi_hip_tree &t, std::vector<uint8_t> marked) {
  const int D = t.D;
  int64_t count = 0;
  for (;;) {
    std::vector<int32_t> r_lv;
    std::vector<int64_t> r_ic;
    for (size_t i = 0; i < t.n(); ++i)
      if (marked[i]) {
        r_lv.push_back(t.lv[i]);
        for (int d = 0; d < D; ++d) r_ic.push_back(t.ic[i * D + d]);
      }
    if (r_lv.empty()) break;
    count += (int64_t)r_lv.size();
    refine_unbalanced(t, marked);
    const CellMap cells(t);
    marked.assign(t.n(), 0);
    int64_t nb[3], p[3], pnb[3];
    for (size_t r = 0; r < r_lv.size(); ++r) {
      const int l = r_lv[r];
      if (l == 0) continue;
      const int64_t *c = &r_ic[r * D];
      for (int dirn = 0; dirn < 2 * D; ++dirn) {
        int32_t v;
        if (shift(t, l, c, dirn, nb) && cells.find(cell_key(l, nb, D), &v)) continue;
        // no same-level cell across this face: refine the parent's neighbour if it is a leaf
        for (int d = 0; d < D; ++d) p[d] = c[d] >> 1;
        if (shift(t, l - 1, p, dirn, pnb) && cells.find(cell_key(l - 1, pnb, D), &v) && v >= 0) marked[v] = 1;
      }
    }
  }
  return count;
}

int64_t coarsen(trixi_hip_tree &t, const std::vector<int32_t> &k_lv, const std::vector<int64_t> &k_ic) {
  const int D = t.D, nch = 1 << D;
  // parents of the candidates that are still leaves (rebalancing may have
  // refined some), each with the number of its children asking
  std::vector<std::pair<uint64_t, size_t>> par;  // (parent key, candidate index)
  {
    const CellMap cells(t);
    int64_t p[3];
    for (size_t i = 0; i < k_lv.size(); ++i) {
      int32_t v;
      if (k_lv[i] <= 0 || !cells.find(cell_key(k_lv[i], &k_ic[i * D], D), &v) || v < 0) continue;
      for (int d = 0; d < D; ++d) p[d] = k_ic[i * D + d] >> 1;
      par.push_back({cell_key(k_lv[i] - 1, p, D), i});
    }
  }
  std::sort(par.begin(), par.end());  // by key; equal keys keep candidate order
  struct Parent {
    uint64_t key;
    int level;
    int64_t c[3];
  };
  std::vector<Parent> full;  // in key order
  for (size_t a = 0; a < par.size();) {
    size_t b = a;
    while (b < par.size() && par[b].first == par[a].first) ++b;
    if ((int)(b - a) == nch) {
      Parent P{par[a].first, k_lv[par[a].second] - 1, {0, 0, 0}};
      for (int d = 0; d < D; ++d) P.c[d] = k_ic[par[a].second * D + d] >> 1;
      full.push_back(P);
    }
    a = b;
  }
  int64_t total = 0;
  int maxl = -1;
  for (const auto &P : full) maxl = std::max(maxl, P.level);
  for (int level = maxl; level >= 0; --level) {
    std::vector<uint64_t> done;
    {
      const CellMap cells(t);
      int64_t ch[3], nb[3];
      for (const auto &P : full) {
        if (P.level != level) continue;
        bool skip = false;
        for (int q = 0; q < nch && !skip; ++q) {
          for (int d = 0; d < D; ++d) ch[d] = 2 * P.c[d] + ((q >> d) & 1);
          for (int dirn = 0; dirn < 2 * D && !skip; ++dirn) {
            const int bit = (q >> (dirn / 2)) & 1;
            if (bit == (dirn % 2 == 0 ? 1 : 0)) continue;  // the neighbour is a sibling
            int32_t v;
            if (shift(t, level + 1, ch, dirn, nb) && cells.find(cell_key(level + 1, nb, D), &v) && v < 0)
              skip = true;
          }
        }
        if (!skip) done.push_back(P.key);
      }
    }
    if (done.empty()) continue;
    total += (int64_t)done.size();
    std::sort(done.begin(), done.end());
    // collapse the children of the coarsened parents (consecutive, child 0 first)
    std::vector<int32_t> lv, kind, src;
    std::vector<int64_t> ic;
    std::vector<double> ctr;
    int64_t p[3];
    for (size_t i = 0; i < t.n(); ++i) {
      bool inpar = false;
      if (t.lv[i] > 0) {
        for (int d = 0; d < D; ++d) p[d] = t.ic[i * D + d] >> 1;
        inpar = std::binary_search(done.begin(), done.end(), cell_key(t.lv[i] - 1, p, D));
      }
      bool first_child = inpar;
      for (int d = 0; d < D && first_child; ++d) first_child = (t.ic[i * D + d] & 1) == 0;
      if (inpar && !first_child) continue;
      if (first_child) {
        const double h = t.length0 / (double)((int64_t)1 << t.lv[i]);
        lv.push_back(t.lv[i] - 1);
        kind.push_back(-1);
        src.push_back(t.src[i]);
        for (int d = 0; d < D; ++d) {
          ic.push_back(t.ic[i * D + d] >> 1);
          ctr.push_back(t.ctr[i * D + d] + h / 2);
        }
      } else {
        lv.push_back(t.lv[i]);
        kind.push_back(t.kind[i]);
        src.push_back(t.src[i]);
        for (int d = 0; d < D; ++d) {
          ic.push_back(t.ic[i * D + d]);
          ctr.push_back(t.ctr[i * D + d]);
        }
      }
    }
    t.lv.swap(lv);
    t.kind.swap(kind);
    t.src.swap(src);
    t.ic.swap(ic);
    t.ctr.swap(ctr);
    t.connected = false;
  }
  return total;
}

// Interfaces, L2 mortars and boundaries of the leaves: a same-level leaf across
// a face is an interface neighbour; a same-level cell with children makes a
// mortar with this element as the large side; no same-level cell means this
// element is a small element of the coarser neighbour's mortar; outside a
// non-periodic domain the face is a boundary.
int connect(trixi_hip_tree &t) {
  const int D = t.D, ns = 1 << (D - 1), F = 2 * D;
  const size_t E = t.n();
  const CellMap cells(t);
  t.nbr.assign(E * F, -1);
  std::vector<uint8_t> boundary(E * F, 0);
  t.m_ids.clear();
  t.m_side.clear();
  t.m_or.clear();
  int64_t nb[3], ch[3];
  for (size_t e = 0; e < E; ++e) {  // mortars in element order, then direction order
    const int l = t.lv[e];
    for (int dirn = 0; dirn < F; ++dirn) {
      if (!shift(t, l, &t.ic[e * D], dirn, nb)) {
        boundary[e * F + dirn] = 1;
        continue;
      }
      int32_t v;
      if (!cells.find(cell_key(l, nb, D), &v)) continue;  // small element of a coarser neighbour's mortar
      if (v >= 0) {
        t.nbr[e * F + dirn] = v;
        continue;
      }
      const int o = dirn / 2, plus = dirn % 2;
      int fr[2] = {-1, -1};
      for (int d = 0, j = 0; d < D; ++d)
        if (d != o) fr[j++] = d;
      for (int q = 0; q < ns; ++q) {
        for (int d = 0; d < D; ++d) ch[d] = 2 * nb[d];
        ch[o] += 1 - plus;
        ch[fr[0]] += q & 1;
        if (D == 3) ch[fr[1]] += q >> 1;
        int32_t w;
        if (!cells.find(cell_key(l + 1, ch, D), &w) || w < 0) {
          t.err = "the tree is not 2:1 balanced";
          return TRIXI_HIP_ESTATE;
        }
        t.m_ids.push_back(w);
      }
      t.m_ids.push_back((int32_t)e);
      t.m_side.push_back(plus ? 1 : 2);
      t.m_or.push_back(o);
    }
  }
  t.if_ids.clear();
  t.if_or.clear();
  for (size_t e = 0; e < E; ++e)
    for (int o = 0; o < D; ++o)
      if (t.nbr[e * F + 2 * o + 1] >= 0) {
        t.if_ids.push_back((int32_t)e);
        t.if_ids.push_back(t.nbr[e * F + 2 * o + 1]);
        t.if_or.push_back(o);
      }
  t.b_ids.clear();
  t.b_or.clear();
  t.b_side.clear();
  for (int dirn = 0; dirn < F; ++dirn)
    for (size_t e = 0; e < E; ++e)
      if (boundary[e * F + dirn]) {
        t.b_ids.push_back((int32_t)e);
        t.b_or.push_back(dirn / 2);
        t.b_side.push_back(dirn % 2 == 1 ? 1 : 2);
      }
  t.connected = true;
  return TRIXI_HIP_OK;
}

}  // namespace

extern "C" {

trixi_hip_tree *trixi_hip_tree_create(int ndims, const int *periodicity, double length_level_0, int nleaves,
                                      const int32_t *levels, const int64_t *coordinates, const double *centers) {
  if ((ndims != 2 && ndims != 3) || !periodicity || !(length_level_0 > 0) || nleaves < 0 ||
      (nleaves > 0 && (!levels || !coordinates || !centers)))
    return nullptr;
  trixi_hip_tree *t = new (std::nothrow) trixi_hip_tree;
  if (!t) return nullptr;
  t->D = ndims;
  for (int d = 0; d < ndims; ++d) t->per[d] = periodicity[d] != 0;
  t->length0 = length_level_0;
  for (int i = 0; i < nleaves; ++i) {
    if (levels[i] < 0 || levels[i] > kMaxLevel) {
      delete t;
      return nullptr;
    }
    for (int d = 0; d < ndims; ++d) {
      const int64_t c = coordinates[(size_t)i * ndims + d];
      if (c < 0 || c >= ((int64_t)1 << levels[i])) {
        delete t;
        return nullptr;
      }
    }
  }
  t->lv.assign(levels, levels + nleaves);
  t->ic.assign(coordinates, coordinates + (size_t)nleaves * ndims);
  t->ctr.assign(centers, centers + (size_t)nleaves * ndims);
  t->kind.assign(nleaves, 0);
  t->src.resize(nleaves);
  for (int i = 0; i < nleaves; ++i) t->src[i] = i;
  return t;
}

void trixi_hip_tree_destroy(trixi_hip_tree *t) { delete t; }

const char *trixi_hip_tree_last_error(const trixi_hip_tree *t) { return t ? t->err.c_str() : "null tree"; }

int trixi_hip_tree_refine(trixi_hip_tree *t, const uint8_t *marked, int64_t *nrefined) {
  if (!t || (!marked && t->n() > 0)) return TRIXI_HIP_EINVAL;
  std::vector<uint8_t> m(marked, marked + t->n());
  for (size_t i = 0; i < t->n(); ++i)
    if (m[i] && t->lv[i] >= kMaxLevel) {
      t->err = "refinement beyond level 19";
      return TRIXI_HIP_EINVAL;
    }
  const int64_t c = refine_balanced(*t, std::move(m));
  for (size_t i = 0; i < t->n(); ++i)
    if (t->lv[i] > kMaxLevel) {
      t->err = "refinement beyond level 19";
      return TRIXI_HIP_EINVAL;
    }
  if (nrefined) *nrefined = c;
  return TRIXI_HIP_OK;
}

int trixi_hip_tree_coarsen(trixi_hip_tree *t, int ncandidates, const int32_t *levels, const int64_t *coordinates,
                           int64_t *ncoarsened) {
  if (!t || ncandidates < 0 || (ncandidates > 0 && (!levels || !coordinates))) return TRIXI_HIP_EINVAL;
  const int D = t->D;
  std::vector<int32_t> lv(levels, levels + ncandidates);
  std::vector<int64_t> ic(coordinates, coordinates + (size_t)ncandidates * D);
  for (int i = 0; i < ncandidates; ++i)
    if (lv[i] < 0 || lv[i] > kMaxLevel) return TRIXI_HIP_EINVAL;
  const int64_t c = coarsen(*t, lv, ic);
  if (ncoarsened) *ncoarsened = c;
  return TRIXI_HIP_OK;
}

int trixi_hip_tree_sizes(trixi_hip_tree *t, int64_t *sizes) {
  if (!t || !sizes) return TRIXI_HIP_EINVAL;
  if (!t->connected) {
    int r = connect(*t);
    if (r) return r;
  }
  sizes[0] = (int64_t)t->n();
  sizes[1] = (int64_t)t->if_or.size();
  sizes[2] = (int64_t)t->m_or.size();
  sizes[3] = (int64_t)t->b_ids.size();
  return TRIXI_HIP_OK;
}

int trixi_hip_tree_leaves(const trixi_hip_tree *t, int32_t *levels, int64_t *coordinates, double *centers,
                          int32_t *kind, int32_t *src) {
  if (!t) return TRIXI_HIP_EINVAL;
  const size_t n = t->n(), D = t->D;
  if (levels) std::memcpy(levels, t->lv.data(), n * 4);
  if (coordinates) std::memcpy(coordinates, t->ic.data(), n * D * 8);
  if (centers) std::memcpy(centers, t->ctr.data(), n * D * 8);
  if (kind) std::memcpy(kind, t->kind.data(), n * 4);
  if (src) std::memcpy(src, t->src.data(), n * 4);
  return TRIXI_HIP_OK;
}

int trixi_hip_tree_connectivity(trixi_hip_tree *t, int32_t *element_neighbors, int32_t *neighbor_ids,
                                int32_t *orientations, int32_t *mortar_neighbor_ids, int32_t *mortar_large_sides,
                                int32_t *mortar_orientations, int32_t *boundary_neighbor_ids,
                                int32_t *boundary_orientations, int32_t *boundary_neighbor_sides) {
  if (!t) return TRIXI_HIP_EINVAL;
  if (!t->connected) {
    int r = connect(*t);
    if (r) return r;
  }
  auto put = [](int32_t *dst, const std::vector<int32_t> &v) {
    if (dst && !v.empty()) std::memcpy(dst, v.data(), v.size() * 4);
  };
  put(element_neighbors, t->nbr);
  put(neighbor_ids, t->if_ids);
  put(orientations, t->if_or);
  put(mortar_neighbor_ids, t->m_ids);
  put(mortar_large_sides, t->m_side);
  put(mortar_orientations, t->m_or);
  put(boundary_neighbor_ids, t->b_ids);
  put(boundary_orientations, t->b_or);
  put(boundary_neighbor_sides, t->b_side);
  return TRIXI_HIP_OK;
}

}  // extern "C"
