// deck_spellings.h -- the Kokkos and Cabana names that deck sources written
// for the reference call on the deck interface, for decks built unchanged
// against deck.h.  Initializers run once on the host before the step loop,
// so every parallel construct here is a serial loop in index order.
//
// Names provided (as used by decks/*.cxx and tests/energy_comparison of the
// reference): KOKKOS_LAMBDA, ExecutionSpace, Kokkos::RangePolicy,
// Kokkos::parallel_for, Cabana::slice (particle members including
// Cell_Index, and field members), Cabana::SimdPolicy,
// Cabana::simd_parallel_for.
#pragma once

#include <cstddef>
#include <type_traits>

#include "pic_types.h"

#ifndef KOKKOS_LAMBDA
#define KOKKOS_LAMBDA [=]
#endif

// the execution space of a deck's initializers: the host, serially
struct HostSerial {};
using ExecutionSpace = HostSerial;

namespace Kokkos {

template <class... Properties> struct RangePolicy {
  std::size_t begin, end;
  RangePolicy(std::size_t b, std::size_t e) : begin(b), end(e) {}
};

// f(i) for i in [begin, end)
template <class... P, class F>
void parallel_for(const char*, const RangePolicy<P...>& policy, F&& f) {
  for (std::size_t i = policy.begin; i < policy.end; ++i) f((int)i);
}
template <class... P, class F> void parallel_for(const RangePolicy<P...>& policy, F&& f) {
  parallel_for("", policy, f);
}
template <class F> void parallel_for(std::size_t n, F&& f, const char* = "") {
  for (std::size_t i = 0; i < n; ++i) f((int)i);
}

}  // namespace Kokkos

namespace Cabana {

template <int Member, class Container> auto slice(Container& c) {
  if constexpr (std::is_same_v<Container, particle_list_t> && Member == Cell_Index)
    return ::cell_slice(c);
  else
    return ::slice<Member>(c);
}

template <int VectorLength, class Space> struct SimdPolicy {
  std::size_t begin, end;
  SimdPolicy(std::size_t b, std::size_t e) : begin(b), end(e) {}
};

// f(s, i) for the particles p in [begin, end), s = p / VectorLength and
// i = p % VectorLength (the AoSoA indices of the reference's containers)
template <int VL, class Space, class F>
void simd_parallel_for(const SimdPolicy<VL, Space>& policy, F&& f, const char* = "") {
  for (std::size_t p = policy.begin; p < policy.end; ++p) f((int)(p / VL), (int)(p % VL));
}

}  // namespace Cabana
