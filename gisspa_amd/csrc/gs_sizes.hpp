// gs_sizes.hpp -- the window sizes L the kernels are compiled for: every
// multiple of 32 from 64 to 1024 (adjustPaddingSize, main.cu:158-178, grows a
// window that is smaller than the templates to the next multiple of 32, and
// the reference's launch grids assume L % 32 == 0) and 720, the window of the
// full-size 5760 x 4092 workload.  Transform lengths L and L/2 factor into
// radices <= 32 (primes up to 31 are direct DFTs, gs_fft.hpp).
#pragma once
#include <type_traits>

#ifndef GS_SIZES
#define GS_SIZES(X) X(64) X(96) X(128) X(160) X(192) X(224) X(256) X(288) X(320) X(352) X(384) \
   X(416) X(448) X(480) X(512) X(544) X(576) X(608) X(640) X(672) X(704) X(736) X(768) X(800) \
   X(832) X(864) X(896) X(928) X(960) X(992) X(1024) X(720)
#endif

namespace gs {

// Calls f(std::integral_constant<int, L>) for a compiled size; false otherwise.
template <class F>
inline bool with_window(int L, F&& f) {
  switch (L) {
#define GS_WINDOW_CASE(n) \
  case n:                 \
    f(std::integral_constant<int, n>{}); \
    return true;
    GS_SIZES(GS_WINDOW_CASE)
#undef GS_WINDOW_CASE
    default:
      return false;
  }
}

}  // namespace gs

// FN<L>(args...) for the run-time size L (nothing for an uncompiled size)
#define GS_DISPATCH(L, FN, ...) \
  ::gs::with_window((L), [&](auto gs_w_) { FN<decltype(gs_w_)::value>(__VA_ARGS__); })
