// pic_types.h -- host-side types of the deck interface.
//
// The input-deck API of the reference (src/input/deck.h) hands particle and
// field containers to user initializers, which fill them through member
// "slices" (src/types.h:31-172).  Here the containers live in host memory
// and are uploaded to the device through the C-ABI afterwards; the member
// names, member order and index helpers are the reference's.
#pragma once

#include <cstddef>
#include <cstdint>
#include <vector>

#ifndef REAL_TYPE
#define real_t float
#else
#define real_t REAL_TYPE
#endif

// Particle members, src/types.h:31-41.
enum UserParticleFields {
  PositionX = 0,
  PositionY,
  PositionZ,
  VelocityX,
  VelocityY,
  VelocityZ,
  Weight,
  Cell_Index,
};

// Field members, src/types.h:138-149.
enum FieldFields {
  FIELD_EX = 0,
  FIELD_EY,
  FIELD_EZ,
  FIELD_CBX,
  FIELD_CBY,
  FIELD_CBZ,
  FIELD_JFX,
  FIELD_JFY,
  FIELD_JFZ
};

// src/input/deck.h:9-12
enum Boundary { Reflect = 0, Periodic };

// Linear voxel index with ghosts, x fastest (src/types.h:195).
#define VOXEL(x, y, z, nx, ny, nz, NG) ((x) + ((nx) + (NG * 2)) * ((y) + ((ny) + (NG * 2)) * (z)))

// Inverse of VOXEL for a row length _x and column height _y (src/types.h:184-193).
#define RANK_TO_INDEX(rank, ix, iy, iz, _x, _y)            \
  do {                                                     \
    const long rank_to_index_r_ = (long)(rank);            \
    const long rank_to_index_q_ = rank_to_index_r_ / (long)(_x); \
    (ix) = rank_to_index_r_ - rank_to_index_q_ * (long)(_x);     \
    (iz) = rank_to_index_q_ / (long)(_y);                  \
    (iy) = rank_to_index_q_ - (iz) * (long)(_y);           \
  } while (0)

// One member of a container, accessed as slice(i) or slice.access(s, i)
// with i = s * vector_length + i (the AoSoA indexing of the reference's
// initializers).
template <class T> class host_slice {
 public:
  host_slice(T* p, std::size_t vl) : p_(p), vl_(vl) {}
  T& operator()(std::size_t i) const { return p_[i]; }
  T& access(std::size_t s, std::size_t i) const { return p_[s * vl_ + i]; }

 private:
  T* p_;
  std::size_t vl_;
};

// Host particle list: SoA of the seven real_t members plus the int cell.
class particle_list_t {
 public:
  static constexpr std::size_t vector_length = 32;

  explicit particle_list_t(std::size_t n = 0) { resize(n); }
  void resize(std::size_t n) {
    n_ = n;
    for (auto& m : real_) m.assign(n, real_t(0));
    cell_.assign(n, 0);
  }
  std::size_t size() const { return n_; }
  // number of vector_length blocks (the "s" range of the reference's SIMD loops)
  std::size_t numSoA() const { return (n_ + vector_length - 1) / vector_length; }
  std::size_t arraySize(std::size_t s) const {
    return (s + 1) * vector_length <= n_ ? vector_length : n_ - s * vector_length;
  }
  real_t* data(int member) { return real_[member].data(); }
  int32_t* cells() { return cell_.data(); }

 private:
  std::size_t n_ = 0;
  std::vector<real_t> real_[7];
  std::vector<int32_t> cell_;
};

template <int Member> host_slice<real_t> slice(particle_list_t& p) {
  static_assert(Member >= PositionX && Member <= Weight, "real_t particle member");
  return host_slice<real_t>(p.data(Member), particle_list_t::vector_length);
}

inline host_slice<int32_t> cell_slice(particle_list_t& p) {
  return host_slice<int32_t>(p.cells(), particle_list_t::vector_length);
}

// The reference's Cabana::simd_parallel_for(policy, f): f(s, i) for every
// particle, here a serial host loop in index order.
template <class F> void simd_for_each(particle_list_t& p, F&& f) {
  for (std::size_t s = 0; s < p.numSoA(); ++s)
    for (std::size_t i = 0; i < p.arraySize(s); ++i) f(s, i);
}

// Host field array: 9 members of Ng = (nx+2)(ny+2)(nz+2) voxels each,
// member-major (the C-ABI's field layout).
class field_array_t {
 public:
  explicit field_array_t(std::size_t ncells = 0) { resize(ncells); }
  void resize(std::size_t ncells) {
    n_ = ncells;
    v_.assign(9 * ncells, real_t(0));
  }
  std::size_t size() const { return n_; }
  real_t* data() { return v_.data(); }
  real_t* member(int m) { return v_.data() + (std::size_t)m * n_; }

 private:
  std::size_t n_ = 0;
  std::vector<real_t> v_;
};

template <int Member> host_slice<real_t> slice(field_array_t& f) {
  static_assert(Member >= FIELD_EX && Member <= FIELD_JFZ, "field member");
  return host_slice<real_t>(f.member(Member), 1);
}
