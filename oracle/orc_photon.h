// ORACLE -- test infrastructure only.
//
// CPU restatement of the reference's photon maps: photons, the
// left-balanced kd-tree and its k-nearest search (photon/photon-map.cc),
// the Gaussian density filter (util/gaussian-filter.h), the photon-
// direction histogram (geometry/dir-hist.h, geometry/dir-hist-dist.h over
// geometry/hist-2d.h and geometry/hist-2d-dist.cc), the radical inverse
// (util/radical-inverse.h) and the power heuristic with sample counts
// (render/mis-sample-weight.h:45-77).
#pragma once
#include <algorithm>
#include <cmath>
#include <vector>

#include "orc_render.h"

namespace orc {

// util/radical-inverse.h:27-43
static inline double radical_inverse(unsigned num, unsigned base) {
  double val = 0;
  double inv_base = 1 / double(base);
  double inv_bi = inv_base;
  while (num > 0) {
    unsigned d_i = num % base;
    val += d_i * inv_bi;
    num /= base;
    inv_bi *= inv_base;
  }
  return val;
}

// render/mis-sample-weight.h:45-77: the power heuristic including the
// 1 / NUM_SAMPLES factor (but 1 for an overflowing term).
static inline float mis_sample_weight(float pdf, float num, float other_pdf, float num_other) {
  float term = num * pdf;
  float other = num_other * other_pdf;
  if (term > 1e10f) return 1;
  if (other > 1e10f) return 0;
  float t2 = term * term, o2 = other * other;
  return (term * pdf) / (t2 + o2);
}

// photon/photon.h:26-53: where the photon arrived, its power, and the
// direction it came from.
template<class R> struct Photon {
  V3<R> pos;
  C3<R> power;
  V3<R> dir;
};

// PhotonMap (photon/photon-map.cc:22-407): the photons of one set as a
// left-balanced kd-tree in heap order (children of node i at 2i+1, 2i+2),
// each node split at its photon on the axis of its range's widest box.
template<class R> struct PhotonMap {
  std::vector<Photon<R>> photons;
  std::vector<unsigned char> axes;

  // photon/photon-map.cc:37-92 (Baerentzen, "On Left-balancing Binary
  // Trees", 2003)
  static unsigned left_child_nodes(unsigned num) {
    unsigned max_bottom = 1;
    while (max_bottom * 2 <= num) max_bottom += max_bottom;
    unsigned non_bottom = max_bottom - 1;
    unsigned bottom = num - non_bottom;
    unsigned left_child_max_bottom = max_bottom / 2;
    unsigned left_child_non_bottom = (non_bottom - 1) / 2;
    unsigned left_child_bottom = std::min(bottom, left_child_max_bottom);
    return left_child_non_bottom + left_child_bottom;
  }

  typedef typename std::vector<Photon<R>>::iterator It;

  // photon/photon-map.cc:112-189
  void make_kdtree(It beg, It end, unsigned target) {
    It median = beg;
    unsigned split_axis = 0;
    if (beg + 1 != end) {
      BBox<R> bb;
      for (It i = beg; i != end; ++i) bb.add(i->pos);
      R max_dim = 0;
      const R dims[3] = {bb.max.x - bb.min.x, bb.max.y - bb.min.y, bb.max.z - bb.min.z};
      for (unsigned a = 0; a < 3; a++)
        if (dims[a] > max_dim) {
          max_dim = dims[a];
          split_axis = a;
        }
      median += left_child_nodes((unsigned)(end - beg));
      const unsigned ax = split_axis;
      std::nth_element(beg, median, end, [ax](const Photon<R> &a, const Photon<R> &b) {
        return coord(a.pos, ax) < coord(b.pos, ax);
      });
      if (median != beg) make_kdtree(beg, median, target * 2 + 1);
      if (median + 1 != end) make_kdtree(median + 1, end, target * 2 + 2);
    }
    photons[target] = *median;
    axes[target] = (unsigned char)split_axis;
  }

  static R coord(const V3<R> &p, unsigned a) { return a == 0 ? p.x : a == 1 ? p.y : p.z; }

  // photon/photon-map.cc:22-33
  void set_photons(std::vector<Photon<R>> &v) {
    photons.resize(v.size());
    axes.resize(v.size());
    if (!photons.empty()) make_kdtree(v.begin(), v.end(), 0);
  }

  // photon/photon-map.cc:214-300: the MAX_PHOTONS photons nearest POS
  // within sqrt(MAX_DIST_SQ), kept in a max-heap on distance; once it is
  // full the search radius shrinks to its farthest photon.  A node is
  // searched after its children, and only when it has both (2i + 2 <
  // n): a last node with a left child alone never has that child
  // searched.
  struct DistCmp {
    const V3<R> *pos;
    bool operator()(const Photon<R> *a, const Photon<R> *b) const {
      return (a->pos - *pos).length_squared() < (b->pos - *pos).length_squared();
    }
  };
  void find_rec(const V3<R> &pos, unsigned node, unsigned k, R &max_dist_sq,
                std::vector<const Photon<R> *> &heap) const {
    const unsigned n = (unsigned)photons.size();
    const Photon<R> &ph = photons[node];
    if (node * 2 + 2 < n) {
      const unsigned a = axes[node];
      const R split_dist = coord(pos, a) - coord(ph.pos, a);
      const unsigned first = node * 2 + ((split_dist < 0) ? 1 : 2);
      const unsigned second = node * 2 + ((split_dist < 0) ? 2 : 1);
      find_rec(pos, first, k, max_dist_sq, heap);
      if (split_dist * split_dist < max_dist_sq) find_rec(pos, second, k, max_dist_sq, heap);
    }
    const R dist_sq = (pos - ph.pos).length_squared();
    if (dist_sq < max_dist_sq) {
      DistCmp cmp{&pos};
      if (heap.size() == k) {
        std::pop_heap(heap.begin(), heap.end(), cmp);
        heap.pop_back();
      }
      heap.push_back(&ph);
      std::push_heap(heap.begin(), heap.end(), cmp);
      if (heap.size() == k) max_dist_sq = (pos - heap.front()->pos).length_squared();
    }
  }
  // photon/photon-map.h:52-61: returns the final search radius squared.
  R find(const V3<R> &pos, unsigned k, R max_dist_sq,
         std::vector<const Photon<R> *> &heap) const {
    if (!photons.empty()) find_rec(pos, 0, k, max_dist_sq, heap);
    return max_dist_sq;
  }
};

// GaussianSqrtFilter (util/gaussian-filter.h:28-97) over squared offsets:
// exp(-alpha d^2) less its value at the radius, rescaled to 1 at 0.
template<class R> struct GaussSqrt {
  R radius_sq, alpha;
  float edge, scale;
  GaussSqrt(R r2, R a)
      : radius_sq(r2), alpha(a / r2), edge(gauss(r2)), scale(1 / (1 - edge)) {}
  float gauss(R d2) const { return float(std::exp(-alpha * d2)); }
  float operator()(R d2) const {
    if (d2 >= radius_sq) return 0;
    return (gauss(d2) - edge) * scale;
  }
};

// DirHist (8 x 8, geometry/dir-hist.h) with Hist2dDist
// (geometry/hist-2d-dist.cc:38-272) and DirHistDist
// (geometry/dir-hist-dist.h): a histogram of directions over the
// sphere-sample parameters, its row and in-row cumulative sums (summed
// in double, stored in float), sampling by lower_bound and the pdf per
// steradian (bin probability x bins / 4 pi).
struct DirHistDist {
  static const unsigned W = 8, H = 8;
  float bins[W * H];
  float row_cum[H], col_cum[W * H];
  void clear() { std::fill(bins, bins + W * H, 0.f); }
  // Hist2d::add (geometry/hist-2d.h:46-51)
  void add(float pu, float pv, float val) {
    unsigned col = std::min(unsigned(pu * W), W - 1);
    unsigned row = std::min(unsigned(pv * H), H - 1);
    bins[col + row * W] += val;
  }
  void calc() {
    double bin_sum = 0;
    for (unsigned i = 0; i < W * H; i++) bin_sum += double(bins[i]);
    double inv = bin_sum == 0 ? 0 : 1 / bin_sum;
    double norm = 0;
    for (unsigned r = 0; r < H; r++) {
      for (unsigned c = 0; c < W; c++) norm += double(bins[r * W + c]) * inv;
      row_cum[r] = float(norm);
    }
    for (unsigned r = 0; r < H; r++) {
      double rs = 0;
      for (unsigned c = 0; c < W; c++) rs += double(bins[r * W + c]);
      double ir = rs == 0 ? 0 : 1 / rs;
      double nr = 0;
      for (unsigned c = 0; c < W; c++) {
        nr += double(bins[r * W + c]) * ir;
        col_cum[r * W + c] = float(nr);
      }
    }
  }
  static unsigned lower(float val, const float *a, unsigned n) {
    return (unsigned)(std::lower_bound(a, a + n, val) - a);
  }
  float bin_pdf(unsigned col, unsigned row) const {
    float rp = row_cum[row];
    if (row != 0) rp -= row_cum[row - 1];
    float cp = col_cum[row * W + col];
    if (col != 0) cp -= col_cum[row * W + col - 1];
    return rp * cp * float(W * H);
  }
  // Hist2dDist::sample (hist-2d-dist.cc:160-218): false when every bin
  // is zero (pdf 0).
  bool sample(float pu, float pv, float &ou, float &ov, float &pdf) const {
    float u = std::min(pu, 1.f), v = std::min(pv, 1.f);
    unsigned row = lower(v, row_cum, H);
    if (row == H) {
      pdf = 0;
      ou = ov = 0;
      return false;
    }
    unsigned col = lower(u, col_cum + row * W, W);
    pdf = bin_pdf(col, row);
    const float cw = 1.f / W, rh = 1.f / H;
    ou = float(col) * cw + std::fmod(pu, cw);
    ov = float(row) * rh + std::fmod(pv, rh);
    return true;
  }
  float pdf(float pu, float pv) const {  // hist-2d-dist.cc:265-273
    int col = clampv(int(pu * float(W)), 0, int(W) - 1);
    int row = clampv(int(pv * float(H)), 0, int(H) - 1);
    return bin_pdf((unsigned)col, (unsigned)row);
  }
};

}  // namespace orc
