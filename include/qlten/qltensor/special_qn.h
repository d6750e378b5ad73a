// Bosonic quantum-number types of the contraction path: U(1), U(1)xU(1),
// Z_n, U(1)xU(1)xZ_n and the trivial representation.  API as the reference's
// special_qn::U1QN (qltensor/special_qn/u1qn.h:27-169), U1U1QN
// (u1u1qn.h:27-189), ZnQN / Z2QN (znqn.h:27-176), U1U1ZnQN
// (u1u1znqn.h:27-227) and TrivialRepQN
// (trivial_rep_qn.h:27-120).  Equality is by hash, as in the reference; the
// hash maps are kept identical so that Index equality (hash based) agrees with
// the reference's.
#ifndef QLTEN_AMD_QLTENSOR_SPECIAL_QN_H
#define QLTEN_AMD_QLTENSOR_SPECIAL_QN_H

#include <cstddef>
#include <iostream>
#include <string>

namespace qlten {
namespace special_qn {

class U1QN {
 public:
  U1QN() : U1QN(0) {}
  U1QN(const int val) : val_(val), hash_(CalcHash(val)) {}  // NOLINT: implicit like the reference
  U1QN(const std::string &, const int val) : U1QN(val) {}

  U1QN operator-() const { return U1QN(-val_); }
  U1QN &operator+=(const U1QN &rhs) {
    val_ += rhs.val_;
    hash_ = CalcHash(val_);
    return *this;
  }
  U1QN operator+(const U1QN &rhs) const { return U1QN(val_ + rhs.val_); }
  U1QN operator-(const U1QN &rhs) const { return U1QN(val_ - rhs.val_); }
  bool operator==(const U1QN &rhs) const { return hash_ == rhs.hash_; }
  bool operator!=(const U1QN &rhs) const { return !(*this == rhs); }

  size_t dim() const { return 1; }
  size_t Hash() const { return hash_; }
  int GetVal() const { return val_; }
  static U1QN Zero() { return U1QN(0); }

  // Stream format (u1qn.h:126-133): the value and the stored hash, one per line.
  void StreamRead(std::istream &is) { is >> val_ >> hash_; }
  void StreamWrite(std::ostream &os) const { os << val_ << "\n" << hash_ << "\n"; }

 private:
  // the value, rotated left by 31 bits (u1qn.h:139-152)
  static size_t CalcHash(int v) {
    const size_t x = static_cast<size_t>(v);
    return (x << 31) | (x >> 33);
  }
  int val_;
  size_t hash_;
};

class U1U1QN {
 public:
  U1U1QN() : U1U1QN(0, 0) {}
  U1U1QN(const int v1, const int v2) : vals_{v1, v2}, hash_(CalcHash(v1, v2)) {}
  U1U1QN(const std::string &, const int v1, const std::string &, const int v2) : U1U1QN(v1, v2) {}

  U1U1QN operator-() const { return U1U1QN(-vals_[0], -vals_[1]); }
  U1U1QN &operator+=(const U1U1QN &rhs) {
    vals_[0] += rhs.vals_[0];
    vals_[1] += rhs.vals_[1];
    hash_ = CalcHash(vals_[0], vals_[1]);
    return *this;
  }
  U1U1QN operator+(const U1U1QN &rhs) const { return U1U1QN(vals_[0] + rhs.vals_[0], vals_[1] + rhs.vals_[1]); }
  U1U1QN operator-(const U1U1QN &rhs) const { return U1U1QN(vals_[0] - rhs.vals_[0], vals_[1] - rhs.vals_[1]); }
  bool operator==(const U1U1QN &rhs) const { return hash_ == rhs.hash_; }
  bool operator!=(const U1U1QN &rhs) const { return !(*this == rhs); }

  size_t dim() const { return 1; }
  size_t Hash() const { return hash_; }
  int GetVal(size_t i) const { return vals_[i]; }
  static U1U1QN Zero() { return U1U1QN(0, 0); }

  // Stream format (u1u1qn.h:131-139): both values and the stored hash, one per line.
  void StreamRead(std::istream &is) { is >> vals_[0] >> vals_[1] >> hash_; }
  void StreamWrite(std::ostream &os) const { os << vals_[0] << "\n" << vals_[1] << "\n" << hash_ << "\n"; }

 private:
  // both values offset by 2^30 and packed as v1 + 2^31 v2, rotated left by
  // 10 bits (u1u1qn.h:151-173)
  static size_t CalcHash(int v1, int v2) {
    const size_t seg = size_t(1) << 30;
    const size_t h = (static_cast<size_t>(v1) + seg) + (static_cast<size_t>(v2) + seg) * (2 * seg);
    return (h << 10) | (h >> 54);
  }
  int vals_[2];
  size_t hash_;
};

// Z_n charge: addition modulo n.  As in the reference, the constructor keeps
// the value it is given, operator+ reduces with %, operator+= subtracts n
// once, and the negation of v != 0 is n - v.
template <size_t n>
class ZnQN {
 public:
  ZnQN() : ZnQN(0) {}
  ZnQN(const int val) : val_(val), hash_(CalcHash(val)) {}  // NOLINT: implicit like the reference

  ZnQN operator-() const { return ZnQN(val_ == 0 ? 0 : static_cast<int>(n) - val_); }
  // the sum taken modulo n as an unsigned number, as in the reference
  ZnQN operator+(const ZnQN &rhs) const { return ZnQN(static_cast<int>(static_cast<size_t>(val_ + rhs.val_) % n)); }
  ZnQN operator-(const ZnQN &rhs) const { return *this + (-rhs); }
  ZnQN &operator+=(const ZnQN &rhs) {
    val_ += rhs.val_;
    if (static_cast<size_t>(val_) >= n) val_ -= static_cast<int>(n);
    hash_ = CalcHash(val_);
    return *this;
  }
  bool operator==(const ZnQN &rhs) const { return hash_ == rhs.hash_; }
  bool operator!=(const ZnQN &rhs) const { return !(*this == rhs); }

  size_t dim() const { return 1; }
  size_t Hash() const { return hash_; }
  int GetVal() const { return val_; }
  static ZnQN Zero() { return ZnQN(0); }

  // Stream format (znqn.h:134-145): the value and the stored hash, one per line.
  void StreamRead(std::istream &is) { is >> val_ >> hash_; }
  void StreamWrite(std::ostream &os) const { os << val_ << "\n" << hash_ << "\n"; }

 private:
  // the vector hash of one item whose hash is the value (znqn.h:158-167)
  static size_t CalcHash(int v) {
    constexpr size_t kP1 = 11400714785074694791ull, kP2 = 14029467366897019727ull,
                     kP5 = 2870177450012600261ull;
    size_t h = kP5 + static_cast<size_t>(v) * kP2;
    h = (h << 31) | (h >> 33);
    h *= kP1;
    return h + (size_t(1) ^ kP5);
  }
  int val_;
  size_t hash_;
};

using Z2QN = ZnQN<2>;

// U(1) x U(1) x Z_n charge (u1u1znqn.h:27-227): both U(1) values add, the
// Z_n value adds as in ZnQN.
template <size_t n>
class U1U1ZnQN {
 public:
  U1U1ZnQN() : U1U1ZnQN(0, 0, 0) {}
  U1U1ZnQN(const int u1, const int u2, const int zn) : u1_{u1, u2}, zn_(zn), hash_(CalcHash(u1, u2, zn)) {}

  U1U1ZnQN operator-() const { return U1U1ZnQN(-u1_[0], -u1_[1], zn_ == 0 ? 0 : static_cast<int>(n) - zn_); }
  U1U1ZnQN &operator+=(const U1U1ZnQN &rhs) {
    u1_[0] += rhs.u1_[0];
    u1_[1] += rhs.u1_[1];
    zn_ += rhs.zn_;
    if (static_cast<size_t>(zn_) >= n) zn_ -= static_cast<int>(n);
    hash_ = CalcHash(u1_[0], u1_[1], zn_);
    return *this;
  }
  U1U1ZnQN operator+(const U1U1ZnQN &rhs) const {
    U1U1ZnQN r(*this);
    r += rhs;
    return r;
  }
  U1U1ZnQN operator-(const U1U1ZnQN &rhs) const { return *this + (-rhs); }
  bool operator==(const U1U1ZnQN &rhs) const { return hash_ == rhs.hash_; }
  bool operator!=(const U1U1ZnQN &rhs) const { return !(*this == rhs); }

  size_t dim() const { return 1; }
  size_t Hash() const { return hash_; }
  int GetVal(size_t i) const { return i < 2 ? u1_[i] : zn_; }
  static U1U1ZnQN Zero() { return U1U1ZnQN(0, 0, 0); }

  // Stream format (u1u1znqn.h:165-179): the three values and the stored hash.
  void StreamRead(std::istream &is) { is >> u1_[0] >> u1_[1] >> zn_ >> hash_; }
  void StreamWrite(std::ostream &os) const {
    os << u1_[0] << "\n" << u1_[1] << "\n" << zn_ << "\n" << hash_ << "\n";
  }

 private:
  // U(1) values offset by 2^20 in bits 0-20 and 21-41, the Z_n value from
  // bit 42, rotated left by 10 bits (u1u1znqn.h:205-225)
  static size_t CalcHash(int u1, int u2, int zn) {
    const size_t seg = size_t(1) << 20;
    const size_t h = (static_cast<size_t>(u1) + seg) + (static_cast<size_t>(u2) + seg) * (2 * seg) +
                     static_cast<size_t>(zn) * (4 * seg * seg);
    return (h << 10) | (h >> 54);
  }
  int u1_[2];
  int zn_;
  size_t hash_;
};

// The trivial representation: one charge, hash 0, nothing in the stream.
class TrivialRepQN {
 public:
  TrivialRepQN operator-() const { return {}; }
  TrivialRepQN &operator+=(const TrivialRepQN &) { return *this; }
  TrivialRepQN operator+(const TrivialRepQN &) const { return {}; }
  TrivialRepQN operator-(const TrivialRepQN &) const { return {}; }
  bool operator==(const TrivialRepQN &) const { return true; }
  bool operator!=(const TrivialRepQN &) const { return false; }
  static size_t dim() { return 1; }
  static size_t Hash() { return 0; }
  static TrivialRepQN Zero() { return {}; }
  void StreamRead(std::istream &) {}
  void StreamWrite(std::ostream &) const {}
};

template <size_t n>
inline std::istream &operator>>(std::istream &is, ZnQN<n> &qn) {
  qn.StreamRead(is);
  return is;
}
template <size_t n>
inline std::ostream &operator<<(std::ostream &os, const ZnQN<n> &qn) {
  qn.StreamWrite(os);
  return os;
}
template <size_t n>
inline std::istream &operator>>(std::istream &is, U1U1ZnQN<n> &qn) {
  qn.StreamRead(is);
  return is;
}
template <size_t n>
inline std::ostream &operator<<(std::ostream &os, const U1U1ZnQN<n> &qn) {
  qn.StreamWrite(os);
  return os;
}
inline std::istream &operator>>(std::istream &is, TrivialRepQN &qn) {
  qn.StreamRead(is);
  return is;
}
inline std::ostream &operator<<(std::ostream &os, const TrivialRepQN &qn) {
  qn.StreamWrite(os);
  return os;
}

inline std::istream &operator>>(std::istream &is, U1QN &qn) {
  qn.StreamRead(is);
  return is;
}
inline std::ostream &operator<<(std::ostream &os, const U1QN &qn) {
  qn.StreamWrite(os);
  return os;
}
inline std::istream &operator>>(std::istream &is, U1U1QN &qn) {
  qn.StreamRead(is);
  return is;
}
inline std::ostream &operator<<(std::ostream &os, const U1U1QN &qn) {
  qn.StreamWrite(os);
  return os;
}

}  // namespace special_qn

using special_qn::TrivialRepQN;
using special_qn::U1QN;
using special_qn::U1U1QN;
using special_qn::U1U1ZnQN;
using special_qn::Z2QN;
using special_qn::ZnQN;

}  // namespace qlten

#endif
