// store.hpp - the per-step archives of the two-sweep algorithms, resident in device memory.
//
// The reference writes every step's eigenvectors (DiagInfo::save, c++/eigen.hpp:389-400, file
// unitaryfn(N)) and density matrices (DensMatElements::save, c++/operators.hpp:84-97, file
// rhofn(N, prefix)) to the work directory and reads them back in the backward sweeps (DiagInfo(N, P,
// remove_files), c++/eigen.hpp:401-419 and 416-418; DensMatElements::load, c++/operators.hpp:98-111;
// called from c++/core.hpp:135,138 and c++/measurements.hpp:205-209). StepStore keeps them in HBM:
// save() takes the step's objects and shares their device buffers (no copy); load() hands them back.
// Past a device-byte budget the earliest saved steps move to pinned host memory (one device-to-host
// copy per step) and load() brings them back into a fresh device arena (one host-to-device copy),
// with the eigenvector views and their U(i) blocks re-pointed. The bytes are counted as the store's:
// a caller that keeps its own references to a saved step keeps that device memory alive too, as the
// reference's in-memory copy would outlive its file.
#pragma once
#include <deque>
#include <limits>
#include <map>
#include <stdexcept>
#include <string>

#include "dmnrg.hpp"

namespace NRG {

// Page-locked host memory (hipHostMalloc), for the spilled steps.
class PinnedBuf {
 public:
  PinnedBuf() = default;
  explicit PinnedBuf(size_t n) : n_(n) {
    if (n) hip_check(hipHostMalloc((void **)&p_, n * sizeof(double), hipHostMallocDefault), "hipHostMalloc");
  }
  ~PinnedBuf() {
    if (p_) (void)hipHostFree(p_);
  }
  PinnedBuf(PinnedBuf &&o) noexcept : p_(o.p_), n_(o.n_) { o.p_ = nullptr; o.n_ = 0; }
  PinnedBuf &operator=(PinnedBuf &&o) noexcept {
    if (this != &o) {
      if (p_) (void)hipHostFree(p_);
      p_ = o.p_; n_ = o.n_; o.p_ = nullptr; o.n_ = 0;
    }
    return *this;
  }
  PinnedBuf(const PinnedBuf &) = delete;
  PinnedBuf &operator=(const PinnedBuf &) = delete;
  double *get() const { return p_; }
  size_t size() const { return n_; }

 private:
  double *p_ = nullptr;
  size_t n_ = 0;
};

class StepStore {
 public:
  explicit StepStore(size_t device_budget_bytes = std::numeric_limits<size_t>::max()) : budget_(device_budget_bytes) {}

  // DiagInfo::save(N, P): replaces an earlier save of the same step
  void save(size_t N, const DiagInfo &diag) {
    drop_diag(N);
    DiagEntry &e = diags_[N];
    e.diag = diag;
    e.bytes = 0;
    for (const auto &[I, eig] : e.diag) e.bytes += 8 * extent(eig.vectors.dev());
    dev_bytes_ += e.bytes;
    order_.push_back(Key{N, ""});
    enforce_budget();
  }
  // DiagInfo(N, P, remove_files): the saved step (sharing the store's device buffers while resident)
  DiagInfo load(size_t N, bool remove = false) {
    auto it = diags_.find(N);
    if (it == diags_.end()) throw std::runtime_error("StepStore: no eigenvectors saved for step " + std::to_string(N));
    DiagInfo out = it->second.spilled ? restore(it->second) : it->second.diag;
    if (remove) drop_diag(N);
    return out;
  }
  bool has(size_t N) const { return diags_.count(N) > 0; }

  // DensMatElements::save(N, P, prefix): the store takes the matrices over
  void save_rho(size_t N, const std::string &prefix, DensMatElements &&rho) {
    check_prefix(prefix);
    const Key k{N, prefix};
    drop_rho(k);
    RhoEntry &e = rhos_[k];
    e.rho = std::move(rho);
    e.bytes = 0;
    for (const auto &[I, m] : e.rho) e.bytes += 8 * m.size1() * m.size2();
    dev_bytes_ += e.bytes;
    order_.push_back(k);
    enforce_budget();
  }
  // DensMatElements::load(N, P, prefix, remove_files): a device copy, or the store's matrices on remove
  DensMatElements load_rho(size_t N, const std::string &prefix, bool remove = false) {
    check_prefix(prefix);
    const Key k{N, prefix};
    auto it = rhos_.find(k);
    if (it == rhos_.end())
      throw std::runtime_error("StepStore: no density matrices '" + prefix + "' saved for step " + std::to_string(N));
    RhoEntry &e = it->second;
    DensMatElements out;
    if (e.spilled) {
      for (const auto &[I, s] : e.shapes) {
        DMatrix m(s.rows, s.cols);
        if (s.rows * s.cols)
          dcopy(m.data(), e.host.get() + s.off, 8 * s.rows * s.cols, hipMemcpyHostToDevice);
        out.emplace(I, std::move(m));
      }
    } else if (remove) {
      out = std::move(e.rho);
    } else {
      for (const auto &[I, m] : e.rho) {
        DMatrix c(m.size1(), m.size2());
        if (m.size1() * m.size2())
          dcopy(c.data(), m.data(), 8 * m.size1() * m.size2(), hipMemcpyDeviceToDevice);
        out.emplace(I, std::move(c));
      }
    }
    if (remove) drop_rho(k);
    return out;
  }
  bool has_rho(size_t N, const std::string &prefix) const { return rhos_.count(Key{N, prefix}) > 0; }

  size_t device_bytes() const { return dev_bytes_; }   // resident in the store
  size_t host_bytes() const { return host_bytes_; }    // spilled to pinned host memory
  size_t budget() const { return budget_; }

 private:
  // the empty prefix names the eigenvectors of a step in the shared key space
  static void check_prefix(const std::string &prefix) {
    if (prefix.empty()) throw std::invalid_argument("StepStore: density matrices need a non-empty prefix");
  }
  struct Key {
    size_t N;
    std::string prefix;   // "" for eigenvectors
    bool operator<(const Key &o) const { return N != o.N ? N < o.N : prefix < o.prefix; }
    bool operator==(const Key &o) const { return N == o.N && prefix == o.prefix; }
  };
  struct Shape {
    size_t off = 0, rows = 0, cols = 0, ld = 0;
  };
  struct DiagEntry {
    DiagInfo diag;   // while spilled: values and metadata only, the views cleared
    bool spilled = false;
    size_t bytes = 0;
    PinnedBuf host;
    std::map<Invar, Shape> vec;                   // eigenvector view of each subspace in `host`
    std::map<Invar, std::vector<Shape>> blocks;   // its U(i) views, off relative to the eigenvectors
  };
  struct RhoEntry {
    DensMatElements rho;
    bool spilled = false;
    size_t bytes = 0;
    PinnedBuf host;
    std::map<Invar, Shape> shapes;
  };

  // elements from the first to the last of a row-major view
  static size_t extent(const DView &v) { return v.empty() ? 0 : (v.rows - 1) * v.ld + v.cols; }

  void enforce_budget() {
    for (auto it = order_.begin(); dev_bytes_ > budget_ && it != order_.end(); ++it) {
      if (it->prefix.empty()) {
        auto d = diags_.find(it->N);
        if (d != diags_.end() && !d->second.spilled) spill(d->second);
      } else {
        auto r = rhos_.find(*it);
        if (r != rhos_.end() && !r->second.spilled) spill(r->second);
      }
    }
  }

  void spill(DiagEntry &e) {
    size_t tot = 0;
    for (const auto &[I, eig] : e.diag) tot += extent(eig.vectors.dev());
    e.host = PinnedBuf(tot);
    size_t off = 0;
    for (auto &[I, eig] : e.diag) {
      const DView v = eig.vectors.dev();
      const size_t n = extent(v);
      if (n) dcopy(e.host.get() + off, v.p, 8 * n, hipMemcpyDeviceToHost);
      e.vec[I] = Shape{off, v.rows, v.cols, v.ld};
      auto &bl = e.blocks[I];
      bl.clear();
      for (size_t b = 0; b < eig.U.size(); b++) {
        const DView u = eig.U.get(b);
        Shape s{0, u.rows, u.cols, u.ld};
        if (!u.empty()) {
          NRG_ASSERT(u.p >= v.p && u.p < v.p + n);
          s.off = (size_t)(u.p - v.p);
        }   // an empty view keeps its shape; it is re-pointed at the eigenvectors' start
        bl.push_back(s);
      }
      eig.vectors.set_device(DView{nullptr, v.rows, v.cols, v.ld}, nullptr);   // releases the store's reference
      eig.U.clear();
      off += n;
    }
    e.spilled = true;
    dev_bytes_ -= e.bytes;
    host_bytes_ += 8 * tot;
  }

  void spill(RhoEntry &e) {
    size_t tot = 0;
    for (const auto &[I, m] : e.rho) tot += m.size1() * m.size2();
    e.host = PinnedBuf(tot);
    size_t off = 0;
    for (const auto &[I, m] : e.rho) {
      const size_t n = m.size1() * m.size2();
      if (n) dcopy(e.host.get() + off, m.data(), 8 * n, hipMemcpyDeviceToHost);
      e.shapes[I] = Shape{off, m.size1(), m.size2(), m.size2()};
      off += n;
    }
    e.rho.clear();   // frees the device matrices
    e.spilled = true;
    dev_bytes_ -= e.bytes;
    host_bytes_ += 8 * tot;
  }

  // a resident copy of a spilled step in one new arena
  static DiagInfo restore(const DiagEntry &e) {
    auto arena = std::make_shared<DBuf<double>>(e.host.size() + 1);
    if (e.host.size())
      dcopy(arena->get(), e.host.get(), 8 * e.host.size(), hipMemcpyHostToDevice);
    DiagInfo out = e.diag;
    for (auto &[I, eig] : out) {
      const Shape &s = e.vec.at(I);
      double *base = arena->get() + s.off;
      eig.vectors.set_device(DView{base, s.rows, s.cols, s.ld}, arena);
      const auto &bl = e.blocks.at(I);
      eig.U.resize(bl.size());
      for (size_t b = 0; b < bl.size(); b++)
        eig.U.set(b, DView{base + bl[b].off, bl[b].rows, bl[b].cols, bl[b].ld});   // empty views keep their shape
    }
    return out;
  }

  void drop_diag(size_t N) {
    auto it = diags_.find(N);
    if (it == diags_.end()) return;
    if (it->second.spilled) host_bytes_ -= 8 * it->second.host.size();
    else dev_bytes_ -= it->second.bytes;
    diags_.erase(it);
    forget(Key{N, ""});
  }
  void drop_rho(const Key &k) {
    auto it = rhos_.find(k);
    if (it == rhos_.end()) return;
    if (it->second.spilled) host_bytes_ -= 8 * it->second.host.size();
    else dev_bytes_ -= it->second.bytes;
    rhos_.erase(it);
    forget(k);
  }
  void forget(const Key &k) {
    for (auto it = order_.begin(); it != order_.end(); ++it)
      if (*it == k) {
        order_.erase(it);
        return;
      }
  }

  size_t budget_;
  size_t dev_bytes_ = 0, host_bytes_ = 0;
  std::map<size_t, DiagEntry> diags_;
  std::map<Key, RhoEntry> rhos_;
  std::deque<Key> order_;   // save order, the earliest first
};

}  // namespace NRG
